// Device runtime and C ABI of librdp_mi.so.
//
// One Context per process owns the HIP stream, the generated blender divider
// table, the staging/input/scratch arenas and the launch plumbing.  A
// Processor is one emulated RDP: its front end (rdp_frontend.cpp) and its
// device-resident RDRAM, hidden RDRAM and TMEM.
//
// Submission groups pending render passes into rounds: round r holds the
// r-th pending pass of every participating processor.  All inputs of all
// rounds go to the device in one copy; each round is then four launches
// (setup, TMEM, bin, raster) whose blocks cover every pass of the round.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/rdp_mi.h"
#include "rdp_digest.hpp"
#include "rdp_dump.hpp"
#include "rdp_frontend.hpp"
#include "rdp_hazard.hpp"
#include "rdp_types.hpp"
#include "rdp_vi.hpp"
#include "rdp_view.hpp"

namespace rdp_mi {

hipError_t launch_setup(const PassDesc *, const uint32_t *, uint32_t, const uint32_t *, uint32_t, hipStream_t);
hipError_t launch_tmem(const PassDesc *, const uint32_t *, uint32_t, const uint32_t *, uint32_t, hipStream_t);
hipError_t launch_bin(const PassDesc *, const uint32_t *, uint32_t, hipStream_t);
hipError_t launch_raster(uint32_t cls, const PassDesc *, const uint32_t *, uint32_t, int32_t, bool, hipStream_t);
hipError_t launch_raster_fill(const PassDesc *, const uint32_t *, uint32_t, int32_t, hipStream_t);
hipError_t launch_raster_round(uint32_t cls, const PassDesc *, const uint32_t *, uint32_t, const uint32_t *, uint32_t,
                               const uint32_t *, uint32_t, uint32_t, hipStream_t);
hipError_t launch_raster_up(uint32_t cls, const PassDesc *, const uint32_t *, uint32_t, int32_t, hipStream_t);
hipError_t launch_raster_fill_up(const PassDesc *, const uint32_t *, uint32_t, int32_t, hipStream_t);
hipError_t launch_domain(const PassDesc *, const uint32_t *, uint32_t, bool post, hipStream_t);
hipError_t launch_ssaa_resolve(const PassDesc *, const uint32_t *, uint32_t, hipStream_t);
hipError_t launch_domain_range(const uint8_t *rdram, const uint8_t *hidden, uint8_t *ref, uint8_t *planes,
                               uint8_t *hplanes, uint32_t rdram_size, uint32_t num_planes, uint32_t addr,
                               uint32_t num_pixels, uint32_t size_log2, hipStream_t s);
hipError_t upload_blender_lut(const uint8_t *lut);

// Launch shape of the raster stage: the XCD chunk of the tile-to-block
// mapping, the lowest occupancy class a pass runs in (passes below it are
// promoted, merging their launches) and the tile a workgroup takes (8x8 with
// one wave, or 16x16 with one wave per 8x8 quadrant).  Chosen per workload by
// the autotuner.
struct LaunchConfig {
	int32_t xcd_chunk = 64;
	uint32_t min_class = 0;
	bool quad = false;        // 16x16 tiles
	bool persistent = false;  // one k_raster_round launch per round (8x8 tiles, classes 0-1)
};

// Workgroups of a persistent raster launch: one-wave workgroups the GPU
// keeps resident at the raster's occupancy (7 waves per SIMD, 4 SIMDs per
// CU); set from the device's CU count at context init.
static uint32_t g_persistent_grid = 256 * 4 * 7;

static thread_local std::string g_last_error;

static int set_error(int code, const char *fmt, ...)
{
	char buf[512];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof(buf), fmt, ap);
	va_end(ap);
	g_last_error = buf;
	return code;
}

#define HIP_CHECK(expr)                                                                                        \
	do {                                                                                                       \
		hipError_t err_ = (expr);                                                                              \
		if (err_ != hipSuccess)                                                                                \
			return set_error(RDP_MI_ERROR_DEVICE, "%s failed: %s", #expr, hipGetErrorString(err_));            \
	} while (0)

static void blender_divider_table(uint8_t *out)
{
	// Bit-serial non-restoring divider (11-bit dividend, 4-bit divisor,
	// 3-bit partial remainder): the reference carries its output as data in
	// luts.hpp:7-1032 and indexes it at blender.h:133-139.
	for (unsigned i = 0; i < 0x8000; i++) {
		unsigned d = (i >> 11) & 0xf, n = i & 0x7ff, nd = (~d) & 0xf;
		unsigned part = (nd + (n >> 8) + 1) & 7, q = 0, prev = 0;
		for (unsigned k = 0; k < 8; k++) {
			unsigned b = (n >> (7 - k)) & 1;
			unsigned t = prev ? nd + (part << 1) + b + 1 : d + (part << 1) + b;
			part = t & 7;
			prev = (t >> 4) & 1;
			q |= prev << (7 - k);
		}
		out[i] = (uint8_t)q;
	}
}

// A device arena that grows on demand.  A fixed arena (memory the library
// does not own, such as part of the static `scratch` array) never grows: a
// request beyond its size fails with RDP_MI_ERROR_LIMIT.
struct DeviceBuffer {
	uint8_t *ptr = nullptr;
	size_t size = 0;
	bool fixed = false;
	int reserve(size_t want, hipStream_t stream)
	{
		if (want <= size) return 0;
		if (fixed) return RDP_MI_ERROR_LIMIT;
		size_t ns = std::max(want, size * 2);
		ns = (ns + 0xfffff) & ~(size_t)0xfffff;
		if (ptr) {
			// Work already queued may still read the old allocation.
			if (hipStreamSynchronize(stream) != hipSuccess) return RDP_MI_ERROR_DEVICE;
			if (hipFree(ptr) != hipSuccess) return RDP_MI_ERROR_DEVICE;
			ptr = nullptr;
			size = 0;
		}
		if (hipMalloc(&ptr, ns) != hipSuccess) return RDP_MI_ERROR_OUT_OF_MEMORY;
		size = ns;
		return 0;
	}
};

struct PinnedBuffer {
	uint8_t *ptr = nullptr;
	size_t size = 0;
	bool fixed = false;
	int reserve(size_t want)
	{
		if (want <= size) return 0;
		if (fixed) return RDP_MI_ERROR_LIMIT;
		size_t ns = std::max(want, size * 2);
		ns = (ns + 0xfffff) & ~(size_t)0xfffff;
		if (ptr) {
			if (hipHostFree(ptr) != hipSuccess) return RDP_MI_ERROR_DEVICE;
			ptr = nullptr;
			size = 0;
		}
		if (hipHostMalloc(&ptr, ns, hipHostMallocDefault) != hipSuccess) return RDP_MI_ERROR_OUT_OF_MEMORY;
		size = ns;
		return 0;
	}
};

struct RoundEvents {
	hipEvent_t fork = nullptr, tmem_done = nullptr, bin_done = nullptr, join = nullptr, prep_done = nullptr;
	hipEvent_t prep_heavy = nullptr;  // setup and binning of the passes above class 0 done
	hipEvent_t pre_raster = nullptr;  // the round's stream just before its raster launches
};

// A submission lane: a stream with its own pinned staging buffer and device
// input/scratch arenas.  Work on one lane is stream ordered, so the arenas
// are reused by consecutive submissions without host waits (except for the
// staging buffer, which waits for the previous upload).
struct Lane {
	hipStream_t stream = nullptr;
	// Parts of a round run on `side` beside `stream`, and setup/binning of
	// all rounds on `prep` ahead of them (events per round).
	hipStream_t side = nullptr, prep = nullptr;
	hipEvent_t start = nullptr;
	std::vector<RoundEvents> events;
	DeviceBuffer input, scratch;
	PinnedBuffer staging;
	size_t last_arena_bytes = 0;  // packed inputs + scratch of the last submission
	hipEvent_t staging_done = nullptr;
	bool staging_busy = false;

	int create()
	{
		HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
		HIP_CHECK(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
		HIP_CHECK(hipStreamCreateWithFlags(&prep, hipStreamNonBlocking));
		HIP_CHECK(hipEventCreateWithFlags(&start, hipEventDisableTiming));
		HIP_CHECK(hipEventCreateWithFlags(&staging_done, hipEventDisableTiming));
		return 0;
	}
	int ensure_round_events(size_t rounds)
	{
		while (events.size() < rounds) {
			RoundEvents e;
			HIP_CHECK(hipEventCreateWithFlags(&e.fork, hipEventDisableTiming));
			HIP_CHECK(hipEventCreateWithFlags(&e.tmem_done, hipEventDisableTiming));
			HIP_CHECK(hipEventCreateWithFlags(&e.bin_done, hipEventDisableTiming));
			HIP_CHECK(hipEventCreateWithFlags(&e.join, hipEventDisableTiming));
			HIP_CHECK(hipEventCreateWithFlags(&e.prep_done, hipEventDisableTiming));
			HIP_CHECK(hipEventCreateWithFlags(&e.prep_heavy, hipEventDisableTiming));
			HIP_CHECK(hipEventCreateWithFlags(&e.pre_raster, hipEventDisableTiming));
			events.push_back(e);
		}
		return 0;
	}
	void release()
	{
		if (stream) (void)hipStreamSynchronize(stream);
		if (side) (void)hipStreamSynchronize(side);
		if (prep) (void)hipStreamSynchronize(prep);
		if (input.ptr) (void)hipFree(input.ptr);
		if (scratch.ptr) (void)hipFree(scratch.ptr);
		if (staging.ptr) (void)hipHostFree(staging.ptr);
		if (staging_done) (void)hipEventDestroy(staging_done);
		for (auto &e : events)
			for (hipEvent_t ev : { e.fork, e.tmem_done, e.bin_done, e.join, e.prep_done, e.prep_heavy, e.pre_raster })
				(void)hipEventDestroy(ev);
		if (start) (void)hipEventDestroy(start);
		if (side) (void)hipStreamDestroy(side);
		if (prep) (void)hipStreamDestroy(prep);
		if (stream) (void)hipStreamDestroy(stream);
		*this = Lane();
	}
};

// ---- allocation-free submission (rdp_mi_batch_flush_noalloc) ----
// Packed inputs and per-pass scratch of that entry point live in `scratch`,
// a static device array every GPU the library runs on holds until the
// process ends: the first kNoAllocInputBytes hold the packed inputs, the rest
// the per-pass scratch.  In host code the array's name is a host-side
// stand-in, not a device address: set-up looks the device address up once
// per device (noalloc_init) and every later use goes through that address.
// The host side of the upload is a static (pageable) array, so the entry
// point allocates neither device nor pinned host memory.
constexpr size_t kNoAllocInputBytes = 16u << 20;
constexpr size_t kNoAllocScratchBytes = 112u << 20;
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) uint8_t scratch[kNoAllocInputBytes + kNoAllocScratchBytes];
static uint8_t *g_scratch_addr[kNoAllocMaxDevices];  // device address of `scratch`, per device
alignas(4096) static uint8_t g_noalloc_staging[kNoAllocInputBytes];

// Once per device, with that device current.
static int noalloc_init(int device)
{
	if (device < 0 || device >= kNoAllocMaxDevices)
		return set_error(RDP_MI_ERROR_UNSUPPORTED, "device %d beyond the %d the library tracks", device, kNoAllocMaxDevices);
	if (g_scratch_addr[device]) return 0;
	void *addr = nullptr;
	HIP_CHECK(hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch)));
	if (!addr) return set_error(RDP_MI_ERROR_DEVICE, "no device address for the static scratch array");
	g_scratch_addr[device] = static_cast<uint8_t *>(addr);
	return 0;
}

struct Context : Lane {
	bool ok = false;
	int device = 0;  // the device current at init; every launch goes there
	uint8_t *d_gamma = nullptr;
	LaunchConfig cfg;  // for immediate submissions (flush, batch flush, serving)
	std::mutex lock;
	uint8_t *d_digest = nullptr;                // rdp_mi_state_digest: one target and its result
	DeviceBuffer view;                          // rdp_mi_framebuffer_view output

	int init()
	{
		if (ok) return 0;
		int n = 0;
		if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return set_error(RDP_MI_ERROR_DEVICE, "no HIP device");
		HIP_CHECK(hipGetDevice(&device));
		int cus = 0;
		if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
			g_persistent_grid = (uint32_t)cus * 4 * 7;
		if (int err = noalloc_init(device)) return err;
		if (int err = create()) return err;
		std::vector<uint8_t> lut(0x8000);
		blender_divider_table(lut.data());
		HIP_CHECK(upload_blender_lut(lut.data()));
		std::vector<uint8_t> gamma(256 + 256 * 64);
		vi_gamma_table(gamma.data());
		HIP_CHECK(hipMalloc(&d_gamma, gamma.size()));
		HIP_CHECK(hipMemcpy(d_gamma, gamma.data(), gamma.size(), hipMemcpyHostToDevice));
		ok = true;
		return 0;
	}
};

static Context &ctx()
{
	static Context c;
	return c;
}

// Command ring of one processor (CommandRing, command_ring.cpp:31-134):
// rdp_mi_enqueue_command copies complete commands in, and the processor's
// ring thread takes everything available at once and decodes it.  Private
// meta commands ride in the same ring so they stay ordered with the RDP
// commands (Op::MetaSignalTimeline and Op::MetaFlush, rdp_common.hpp:41-42).
constexpr uint32_t kMetaSignal = 1, kMetaFlush = 2;
constexpr size_t kRingWords = 1u << 16;

struct CommandRing {
	std::vector<uint32_t> words = std::vector<uint32_t>(kRingWords);
	uint64_t write_count = 0, read_count = 0, completed_count = 0;
	std::mutex lock;
	std::condition_variable cv;  // data for the thread, space for producers, completion for drains
	std::thread thr;
	bool stop = false;
	bool pending = false;  // the processor held undelivered passes after the last batch
};

struct Processor {
	Frontend fe;
	uint32_t rdram_size = 0;
	uint8_t *d_rdram = nullptr;
	uint8_t *d_hidden = nullptr;
	uint16_t *d_tmem = nullptr;
	uint8_t *user_rdram = nullptr;
	std::vector<uint8_t> rdram_mirror, hidden_mirror;
	uint16_t tmem_mirror[2048];
	uint32_t vi_regs[RDP_MI_VI_COUNT] = {};
	ViState vi;
	rdp_mi_validation_fn validation_fn = nullptr;
	void *validation_user = nullptr;
	bool supported = false;
	// rdp_mi_set_hold_submissions: completed passes stay on the host until an
	// explicit flush, a recording or a batched submission takes them.
	std::atomic<bool> hold{ false };

	// Guards the front end and the VI state: held by the ring thread while
	// it decodes and by every entry point that touches them directly.
	std::mutex lock;
	std::unique_ptr<CommandRing> ring;  // created by the first rdp_mi_enqueue_command
	std::mutex ring_create;

	// Timeline (rdp_device.cpp:1053-1080, 1184-1188): `timeline` counts the
	// values handed out by signal_timeline; `timeline_done` is the last value
	// whose work the completion thread has seen finish.
	std::atomic<uint64_t> timeline{ 0 };
	std::mutex tl_lock;
	std::condition_variable tl_cv;
	uint64_t timeline_done = 0;
	int timeline_error = 0;
	std::string timeline_error_msg;
	// The first failed submission of the processor and its message.
	int submit_error = 0;
	std::string submit_error_msg;

	// RDPDUMP2 capture of everything sent through rdp_mi_enqueue_command,
	// the VI registers and scanouts (rdp_device.cpp:60-73, 874-897, 970-975,
	// 1087-1092).  dump_in_list: the RDRAM state at the start of the current
	// command list is already in the capture.
	std::unique_ptr<DumpWriter> dump;
	bool dump_in_list = false;
	std::mutex dump_lock;

	// Host-coherent RDRAM (created with the caller's RDRAM pointer; the
	// reference's incoherent host-memory path, rdp_renderer.cpp:2523-2907,
	// masked_rdram_resolve.comp:47-70): the caller's memory is RDRAM and the
	// device copy a cache.  Before a submission the pages its passes read
	// (framebuffer rows, TMEM upload sources) are copied in from the caller;
	// the raster marks every byte it writes in d_wmask; after the submission
	// the written pages and their mask are copied out and, once the work is
	// complete, only the marked bytes are merged into the caller's memory, so
	// bytes the caller wrote meanwhile elsewhere in those pages survive.
	bool coherent = false;
	uint8_t *d_wmask = nullptr;
	PinnedBuffer wb_data, wb_mask;
	std::vector<ByteRange> wb_runs;  // page runs of the write-back in flight
	hipEvent_t wb_event = nullptr;
	bool wb_pending = false;
	std::mutex wb_lock;

	// Upscaled rendering (RDP_MI_FLAG_UPSCALING_*; the reference's
	// init_internal_upscaling_factor, rdp_renderer.cpp:545-601): up_log2 > 0
	// holds 4^up_log2 sample planes of RDRAM and hidden RDRAM and the
	// reference copy of RDRAM the domain last agreed with.
	uint32_t up_log2 = 0;
	uint8_t *d_up_rdram = nullptr, *d_up_hidden = nullptr, *d_up_ref = nullptr;
	// Super-sampled readback (RDP_MI_FLAG_SUPER_SAMPLED_READ_BACK): a write
	// mask word per native pixel of a pass's region (1024 x 1024 at most).
	bool ssaa = false, ssaa_dither = false;
	uint32_t *d_up_wmask = nullptr;

	explicit Processor(uint32_t size) : fe(size), rdram_size(size) { memset(tmem_mirror, 0, sizeof(tmem_mirror)); }
	uint8_t *host_rdram() { return user_rdram ? user_rdram : rdram_mirror.data(); }
};

static size_t align_up(size_t v, size_t a) { return (v + a - 1) & ~(a - 1); }

struct PassRef {
	Processor *proc;
	HostPass *pass;
};

// Byte sizes of the per-pass device scratch.
struct ScratchLayout {
	size_t setups, spans, tilerows, bins, tmem, tmem_blocks, psetups, tilerow_prim, total;
};

// Bin groups of a pass: one per 64 primitives, padded to a multiple of 8 so
// the binning stage writes each tile's masks as whole 64-byte records.
static uint32_t bin_groups(size_t num_prims)
{
	return (((uint32_t)num_prims + 63) / 64 + 7) & ~7u;
}

static ScratchLayout scratch_layout(const HostPass &hp)
{
	ScratchLayout l;
	const uint32_t tx = (hp.fb_width + 7) / 8, ty = (hp.fb_height + 7) / 8;
	const uint32_t groups = bin_groups(hp.prims.size());
	l.setups = align_up(hp.prims.size() * sizeof(RasterPrim), 256);
	l.spans = align_up((size_t)hp.num_spans * sizeof(SpanSetup), 256);
	l.tilerows = align_up((size_t)hp.num_tilerows * 4, 256);
	l.bins = align_up((size_t)tx * ty * groups * 8, 256);
	l.tmem = align_up((hp.uploads.size() + 1) * 4096, 256);
	l.tmem_blocks = align_up(std::max<size_t>((hp.uploads.size() + 31) / 32, 1) * 2048 * 4, 256);
	l.psetups = align_up(hp.prims.size() * sizeof(PrimSetup), 256);
	l.tilerow_prim = align_up((size_t)hp.num_tilerows * 2, 256);
	l.total = l.setups + l.spans + l.tilerows + l.bins + l.tmem + l.tmem_blocks + l.psetups + l.tilerow_prim;
	return l;
}

// One round: the four stage launches covering every pass of the round (the
// raster stage as one launch per occupancy class present).
struct Launch {
	const PassDesc *descs;
	const uint32_t *items[4];  // setup, TMEM gather, bin, raster
	uint32_t n[4];
	const uint32_t *tmem_scan_items;
	uint32_t tmem_scan_n;
	const uint32_t *span_items;  // second setup launch: one item per 8 tile-row entries
	uint32_t span_n;
	// The setup, span and bin items of passes above occupancy class 0 come
	// first in their lists: run ahead on their own, they let the heavier
	// classes' raster start while the light passes are still being set up.
	uint32_t setup_hi, span_hi, bin_hi;
	const uint32_t *raster_items[kRasterClasses];  // 8x8 tiles
	uint32_t raster_n[kRasterClasses];
	const uint32_t *quad_items[kRasterClasses];  // the same passes as 16x16 tiles
	uint32_t quad_n[kRasterClasses];
	const uint32_t *fill_items;  // passes of FILL primitives only: k_raster_fill, 4 tiles per item
	uint32_t fill_n;
	// Upscaled rendering: the domain's pre updates (before the raster) and
	// post updates (after it) of the round's native passes, and the 8x8 tiles
	// of its upscaled passes by class (after the post updates).
	const uint32_t *domain_pre_items, *domain_post_items;
	uint32_t domain_pre_n, domain_post_n;
	const uint32_t *up_items[kRasterClasses];
	uint32_t up_n[kRasterClasses];
	const uint32_t *up_fill_items;  // upscaled passes of FILL primitives only: k_raster_fill_up
	uint32_t up_fill_n;
	const uint32_t *resolve_items;  // the round's upscaled passes with a resolve, after their raster
	uint32_t resolve_n;
	bool prep_ahead;  // every round has its own scratch: setup/binning may run ahead
	bool tmem_early;  // no upload reads what the previous round's raster writes
	uint64_t raster_bytes;  // algorithmic framebuffer bytes read + written by k_raster
	uint64_t class_bytes[kRasterClasses], fill_bytes;  // the same per raster launch
	uint32_t fill_tiles;
	uint32_t passes;
	// Raster work items of class c in the configured tile size.
	const uint32_t *ritems(uint32_t c, bool quad) const { return quad ? quad_items[c] : raster_items[c]; }
	uint32_t rn(uint32_t c, bool quad) const { return quad ? quad_n[c] : raster_n[c]; }
};

// Events around each raster launch of a round when the stages run in series
// (profiling): begin and end of k_raster_fill, then of each class launch.
constexpr size_t kRasterEventsPerRound = 2 * (1 + kRasterClasses);

// Whether round r's TMEM replay may start before round r-1's raster: no
// upload of a pass reads RDRAM that the same processor's pass in round r-1
// can write.
static bool tmem_may_start_early(const std::vector<std::vector<PassRef>> &rounds, size_t r)
{
	if (r == 0) return false;
	std::vector<ByteRange> writes;
	for (const PassRef &cur : rounds[r]) {
		if (cur.pass->uploads.empty()) continue;
		const HostPass *prev = nullptr;
		for (const PassRef &p : rounds[r - 1])
			if (p.proc == cur.proc) prev = p.pass;
		if (!prev) continue;
		writes.clear();
		pass_write_ranges(*prev, writes);
		for (const UploadInfo &u : cur.pass->uploads) {
			ByteRange rd;
			if (!upload_read_range(u, rd)) return false;
			for (const ByteRange &w : writes)
				if (ranges_overlap(rd, w, cur.proc->rdram_size)) return false;
		}
	}
	return true;
}

// Blocks of 256 elements of a native pass's domain update: the framebuffer
// region fb_width x domain_height in whole words (rdp_renderer.cpp:1966-1973).
static uint32_t domain_blocks(const HostPass &hp)
{
	const uint32_t size_log2 = hp.fb_fmt == FB_RGBA8888 ? 2 : (hp.fb_fmt == FB_RGBA5551 || hp.fb_fmt == FB_IA88) ? 1 : 0;
	const uint32_t align = 4u >> size_log2;
	const uint32_t n = (hp.fb_width * hp.domain_height + align - 1u) & ~(align - 1u);
	return (n + 255u) / 256u;
}

// Blocks of 256 native pixels of an upscaled pass's resolve region.
static uint32_t resolve_blocks(const HostPass &hp)
{
	return ((hp.fb_width >> hp.scale_log2) * hp.domain_height + 255u) / 256u;
}

// A pass whose every primitive is a FILL-mode primitive.
static bool fill_only(const HostPass &hp)
{
	for (const PrimRecord &p : hp.prims)
		if (p.variant != kFillVariant) return false;
	return !hp.prims.empty();
}

static uint64_t framebuffer_bytes(const HostPass &hp)
{
	// Colour value + hidden bits, and 16-bit depth + hidden bits, per pixel.
	uint64_t color;
	switch (hp.fb_fmt) {
	case FB_I4:
	case FB_I8: color = 2; break;              // 1 byte + 1 hidden byte per 2 pixels, rounded up
	case FB_RGBA8888: color = 6; break;        // 4 bytes + 2 hidden bytes
	default: color = 3; break;                 // 2 bytes + 1 hidden byte
	}
	const uint64_t depth = hp.fb_addr == hp.depth_addr ? 0 : 3;
	const uint64_t tiles = (uint64_t)((hp.fb_width + 7) / 8) * ((hp.fb_height + 7) / 8);
	if (hp.prims.empty()) return 0;
	// A fill-only pass writes colour and reads nothing.
	if (fill_only(hp)) return tiles * 64 * color;
	return 2 * tiles * 64 * (color + depth);
}

// Launches every round on the lane's stream.  In submissions with a round of
// many passes, three streams share the work:
//   - `prep` runs the setup and binning of every round ahead, starting with
//     the submission (they read only the packed inputs, and each round has
//     its own scratch region), so round r+1's setup overlaps round r's
//     raster;
//   - `side` runs a round's TMEM replay, which depends on RDRAM as the
//     previous round left it and writes only its own snapshots, beside the
//     rest of the round, and then the raster launches of the heavier
//     occupancy classes beside the light class's launch on the main stream,
//     so the light launch fills the GPU while the heavy one's long per-tile
//     chains finish;
//   - the main stream joins everything back before the next round's TMEM.
// Rounds of a few passes (the plain path, small serving batches) stay in
// series, where the events cost more than they hide; stage timing
// (stage_events) needs the stages in series too.
constexpr uint32_t kOverlapMinPasses = 8;

static int execute_launches(const std::vector<Launch> &launches, Lane &lane, const LaunchConfig &cfg,
                            hipEvent_t *stage_events = nullptr, hipEvent_t *raster_events = nullptr)
{
	hipStream_t s = lane.stream, side = lane.side, prep = lane.prep;
	const bool overlap = !stage_events && side && prep;
	const bool ahead = overlap && !launches.empty() && launches[0].prep_ahead;
	if (overlap)
		if (int err = lane.ensure_round_events(launches.size())) return err;
	if (ahead) {
		HIP_CHECK(hipEventRecord(lane.start, s));
		HIP_CHECK(hipStreamWaitEvent(prep, lane.start, 0));
		for (size_t r = 0; r < launches.size(); r++) {
			const Launch &L = launches[r];
			// The heavier classes' passes first (prep_heavy), then the rest.
			HIP_CHECK(launch_setup(L.descs, L.items[0], L.setup_hi, L.span_items, L.span_hi, prep));
			HIP_CHECK(launch_bin(L.descs, L.items[2], L.bin_hi, prep));
			HIP_CHECK(hipEventRecord(lane.events[r].prep_heavy, prep));
			HIP_CHECK(launch_setup(L.descs, L.items[0] + L.setup_hi, L.n[0] - L.setup_hi, L.span_items + L.span_hi,
			                       L.span_n - L.span_hi, prep));
			HIP_CHECK(launch_bin(L.descs, L.items[2] + L.bin_hi, L.n[2] - L.bin_hi, prep));
			HIP_CHECK(hipEventRecord(lane.events[r].prep_done, prep));
		}
	}
	bool prev_tmem_forked = false;
	for (size_t r = 0; r < launches.size(); r++) {
		const Launch &L = launches[r];
		const bool big = overlap && L.passes >= kOverlapMinPasses;
		const bool fork_tmem = big && L.tmem_scan_n;
		// Raster launches after the class floor: the merged light launch and
		// the heavier classes.
		const uint32_t lo = std::min(cfg.min_class, kRasterClasses - 1);
		uint32_t merged = 0, heavy = 0;
		for (uint32_t c = 0; c <= lo; c++) merged += L.rn(c, cfg.quad);
		for (uint32_t c = lo + 1; c < kRasterClasses; c++) heavy += L.rn(c, cfg.quad);
		// One persistent launch for the round's raster work (8x8 tiles), unless
		// the round has top-class passes (the generic code keeps its own
		// register budget) or the launches are being timed one by one.
		const bool persist = cfg.persistent && !raster_events && L.raster_n[kRasterClasses - 1] == 0 &&
		                     L.raster_n[0] + L.raster_n[1] + L.fill_n > 0;
		const bool fork_raster = big && merged && heavy && !persist;
		const RoundEvents E = overlap ? lane.events[r] : RoundEvents();

		const bool early = fork_tmem && L.tmem_early && r > 0;
		if (fork_tmem) {
			// The replay reads RDRAM as the previous round left it, unless
			// none of its uploads reads what that round's raster writes:
			// then it starts before that raster (`early`).  E.fork marks the
			// end of the previous round's raster launches on the main stream.
			// In round 1 of a submission whose setup runs ahead, an early
			// replay needs only what precedes the submission (lane.start)
			// when round 0's own TMEM replay is ahead of it on the side
			// stream.
			HIP_CHECK(hipEventRecord(E.fork, s));
			hipEvent_t after = E.fork;
			if (early) after = (r == 1 && ahead && prev_tmem_forked) ? lane.start : lane.events[r - 1].pre_raster;
			HIP_CHECK(hipStreamWaitEvent(side, after, 0));
			HIP_CHECK(launch_tmem(L.descs, L.items[1], L.n[1], L.tmem_scan_items, L.tmem_scan_n, side));
			HIP_CHECK(hipEventRecord(E.tmem_done, side));
		}
		prev_tmem_forked = fork_tmem || !L.tmem_scan_n;
		if (ahead) {
			if (!fork_tmem) HIP_CHECK(launch_tmem(L.descs, L.items[1], L.n[1], L.tmem_scan_items, L.tmem_scan_n, s));
			HIP_CHECK(hipStreamWaitEvent(s, E.prep_done, 0));
		} else {
			HIP_CHECK(launch_setup(L.descs, L.items[0], L.n[0], L.span_items, L.span_n, s));
			if (stage_events) HIP_CHECK(hipEventRecord(stage_events[r * 4 + 0], s));
			if (!fork_tmem) HIP_CHECK(launch_tmem(L.descs, L.items[1], L.n[1], L.tmem_scan_items, L.tmem_scan_n, s));
			if (stage_events) HIP_CHECK(hipEventRecord(stage_events[r * 4 + 1], s));
			HIP_CHECK(launch_bin(L.descs, L.items[2], L.n[2], s));
			if (stage_events) HIP_CHECK(hipEventRecord(stage_events[r * 4 + 2], s));
		}
		if (overlap) HIP_CHECK(hipEventRecord(E.pre_raster, s));
		// The domain's pre update reads native RDRAM before this round's
		// raster writes it (submit_render_pass, rdp_renderer.cpp:2240-2246).
		HIP_CHECK(launch_domain(L.descs, L.domain_pre_items, L.domain_pre_n, false, s));
		if (persist) {
			if (fork_tmem) HIP_CHECK(hipStreamWaitEvent(s, E.tmem_done, 0));
			// Class 1's tiles first (the longest chains), class 0's, then the fills.
			const uint32_t cls = (L.raster_n[1] || cfg.min_class >= 1) ? 1 : 0;
			const uint32_t total = L.raster_n[0] + L.raster_n[1] + L.fill_n;
			HIP_CHECK(launch_raster_round(cls, L.descs, L.raster_items[1], L.raster_n[1], L.raster_items[0], L.raster_n[0],
			                              L.fill_items, L.fill_n, std::min(total, g_persistent_grid), s));
		} else if (fork_raster) {
			// The heavy launch needs setup, binning and the TMEM replay.
			// With both forked, prep_heavy covers the first two (the heavy
			// launch's classes are all above class 0) and the side stream's
			// own order the third; otherwise the main stream has reached all
			// of them here.
			// The domain's pre update (main stream) must read native RDRAM
			// before any raster of the round writes it.
			hipEvent_t ready = E.prep_heavy;
			if (!(ahead && fork_tmem) || L.domain_pre_n) {
				HIP_CHECK(hipEventRecord(E.bin_done, s));
				ready = E.bin_done;
			}
			HIP_CHECK(hipStreamWaitEvent(side, ready, 0));
			// After an early replay, side-stream order no longer puts the
			// previous round's raster first.
			if (early) HIP_CHECK(hipStreamWaitEvent(side, E.fork, 0));
			for (uint32_t c = lo + 1; c < kRasterClasses; c++)
				HIP_CHECK(launch_raster(c, L.descs, L.ritems(c, cfg.quad), L.rn(c, cfg.quad), cfg.xcd_chunk, cfg.quad, side));
			HIP_CHECK(hipEventRecord(E.join, side));
			if (fork_tmem) HIP_CHECK(hipStreamWaitEvent(s, E.tmem_done, 0));
			HIP_CHECK(launch_raster_fill(L.descs, L.fill_items, L.fill_n, cfg.xcd_chunk, s));
			HIP_CHECK(launch_raster(lo, L.descs, L.ritems(0, cfg.quad), merged, cfg.xcd_chunk, cfg.quad, s));
			HIP_CHECK(hipStreamWaitEvent(s, E.join, 0));
		} else {
			if (fork_tmem) HIP_CHECK(hipStreamWaitEvent(s, E.tmem_done, 0));
			hipEvent_t *re = raster_events ? raster_events + r * kRasterEventsPerRound : nullptr;
			if (re) HIP_CHECK(hipEventRecord(re[0], s));
			HIP_CHECK(launch_raster_fill(L.descs, L.fill_items, L.fill_n, cfg.xcd_chunk, s));
			if (re) HIP_CHECK(hipEventRecord(re[1], s));
			// Class segments are contiguous in class order, so promoting the
			// classes below min_class merges them into one launch.
			if (re) HIP_CHECK(hipEventRecord(re[2 + 2 * lo], s));
			HIP_CHECK(launch_raster(lo, L.descs, L.ritems(0, cfg.quad), merged, cfg.xcd_chunk, cfg.quad, s));
			if (re) HIP_CHECK(hipEventRecord(re[3 + 2 * lo], s));
			for (uint32_t c = lo + 1; c < kRasterClasses; c++) {
				if (re) HIP_CHECK(hipEventRecord(re[2 + 2 * c], s));
				HIP_CHECK(launch_raster(c, L.descs, L.ritems(c, cfg.quad), L.rn(c, cfg.quad), cfg.xcd_chunk, cfg.quad, s));
				if (re) HIP_CHECK(hipEventRecord(re[3 + 2 * c], s));
			}
		}
		// After the native raster: the post updates, then the upscaled passes
		// (rdp_renderer.cpp:2269-2278, 2287-2342, in flush_queues' order).
		HIP_CHECK(launch_domain(L.descs, L.domain_post_items, L.domain_post_n, true, s));
		HIP_CHECK(launch_raster_fill_up(L.descs, L.up_fill_items, L.up_fill_n, cfg.xcd_chunk, s));
		for (uint32_t c = 0; c < kRasterClasses; c++)
			HIP_CHECK(launch_raster_up(c, L.descs, L.up_items[c], L.up_n[c], cfg.xcd_chunk, s));
		HIP_CHECK(launch_ssaa_resolve(L.descs, L.resolve_items, L.resolve_n, s));
		if (stage_events) HIP_CHECK(hipEventRecord(stage_events[r * 4 + 3], s));
	}
	return 0;
}

// Pack the inputs of all rounds into `input` (one async copy on the context
// stream) and describe the launches.  Caller holds the context lock.
static int prepare_rounds(Lane &c, const std::vector<std::vector<PassRef>> &rounds, DeviceBuffer &input,
                          DeviceBuffer &scratch, std::vector<Launch> &launches, PinnedBuffer *staging_override = nullptr)
{
	PinnedBuffer &staging = staging_override ? *staging_override : c.staging;
	launches.clear();
	if (rounds.empty()) return 0;
	if (c.staging_busy) {
		HIP_CHECK(hipEventSynchronize(c.staging_done));
		c.staging_busy = false;
	}

	// Input layout: per round [PassDesc x n][items: setup|tmem|bin|raster], then per pass data.
	struct RoundLayout {
		size_t descs, items_setup, items_tmem, items_bin, items_raster, items_domain, items_up;
		uint32_t n_setup, n_span, n_tmem, n_tmem_gather, n_bin, n_raster, n_fill, n_quad;
		uint32_t n_raster_class[kRasterClasses], n_quad_class[kRasterClasses];
		uint32_t n_dom_pre, n_dom_post, n_up, n_up_class[kRasterClasses], n_up_fill, n_resolve;
		size_t scratch;
	};
	std::vector<RoundLayout> rl(rounds.size());
	struct PassInputs {
		size_t words, states, prims, uploads;
		ScratchLayout sl;
		size_t scratch_base;
		bool fill;  // FILL primitives only (k_raster_fill)
	};
	std::vector<std::vector<PassInputs>> pl(rounds.size());

	// Submissions with a round of many passes run every round's setup and
	// binning ahead on the lane's prep stream (execute_launches), so each
	// round keeps its own scratch region; otherwise rounds run in series and
	// share one region.
	// An upscaled pass reads the TMEM snapshots its native pass (the round
	// before) left in that round's scratch, so upscaled submissions keep
	// every round's scratch too.
	bool prep_ahead = false;
	for (const auto &round : rounds) {
		prep_ahead = prep_ahead || round.size() >= kOverlapMinPasses;
		for (const PassRef &pr : round) prep_ahead = prep_ahead || pr.pass->scale_log2 != 0;
	}
	size_t off = 0, max_scratch = 0, scratch_base = 0;
	for (size_t r = 0; r < rounds.size(); r++) {
		// Work items carry the pass index above 17 bits of per-pass index.
		if (rounds[r].size() > (1u << 15))
			return set_error(RDP_MI_ERROR_UNSUPPORTED, "%zu passes in one round (at most 32768)", rounds[r].size());
		RoundLayout &L = rl[r];
		L.n_setup = L.n_span = L.n_tmem = L.n_tmem_gather = L.n_bin = L.n_raster = L.n_fill = L.n_quad = 0;
		L.n_dom_pre = L.n_dom_post = L.n_up = L.n_up_fill = L.n_resolve = 0;
		for (uint32_t c = 0; c < kRasterClasses; c++) L.n_raster_class[c] = L.n_quad_class[c] = L.n_up_class[c] = 0;
		size_t scr = scratch_base;
		pl[r].resize(rounds[r].size());
		for (size_t i = 0; i < rounds[r].size(); i++) {
			const HostPass &hp = *rounds[r][i].pass;
			const uint32_t tx = (hp.fb_width + 7) / 8, ty = (hp.fb_height + 7) / 8;
			const uint32_t groups = bin_groups(hp.prims.size());
			const bool up = hp.scale_log2 != 0;
			if (up && (i >= (1u << 12) || tx > 1024 || ty > 1024))
				return set_error(RDP_MI_ERROR_UNSUPPORTED, "upscaled round of %zu passes, %ux%u tiles (at most 4096 "
				                 "passes of 1024x1024 tiles)", rounds[r].size(), tx, ty);
			L.n_setup += (uint32_t)(hp.prims.size() + 63) / 64;
			L.n_span += (hp.num_tilerows + 7) / 8;
			// An upscaled pass replays no uploads: it reads its native pass's
			// TMEM snapshots.
			if (!up) {
				L.n_tmem += 8;
				L.n_tmem_gather += 8 * (uint32_t)((hp.uploads.size() + 31) / 32);
			}
			if (hp.domain_pre) L.n_dom_pre += domain_blocks(hp);
			if (hp.domain_post) L.n_dom_post += domain_blocks(hp);
			if (hp.ssaa_resolve) L.n_resolve += resolve_blocks(hp);
			PassInputs &pi = pl[r][i];
			pi.fill = fill_only(hp);
			if (!hp.prims.empty()) {
				L.n_bin += ty * (groups / 8);
				if (up && pi.fill) {
					L.n_up_fill += ty * ((tx + kFillTilesPerItem - 1) / kFillTilesPerItem);
				} else if (up) {
					L.n_up += tx * ty;
					L.n_up_class[std::min(hp.raster_class, kRasterClasses - 1)] += tx * ty;
				} else if (pi.fill) {
					L.n_fill += ty * ((tx + kFillTilesPerItem - 1) / kFillTilesPerItem);
				} else {
					L.n_raster += tx * ty;
					L.n_raster_class[std::min(hp.raster_class, kRasterClasses - 1)] += tx * ty;
					const uint32_t quads = ((tx + 1) / 2) * ((ty + 1) / 2);
					L.n_quad += quads;
					L.n_quad_class[std::min(hp.raster_class, kRasterClasses - 1)] += quads;
				}
			}
			pi.sl = scratch_layout(hp);
			pi.scratch_base = scr;
			scr += pi.sl.total;
		}
		max_scratch = std::max(max_scratch, scr);
		if (prep_ahead) scratch_base = align_up(scr, 256);
		L.scratch = scr;
		L.descs = off;
		off = align_up(off + rounds[r].size() * sizeof(PassDesc), 256);
		L.items_setup = off;
		off = align_up(off + (L.n_setup + L.n_span) * 4ull, 256);
		L.items_tmem = off;
		off = align_up(off + (L.n_tmem + L.n_tmem_gather) * 4ull, 256);
		L.items_bin = off;
		off = align_up(off + L.n_bin * 4ull, 256);
		L.items_raster = off;
		off = align_up(off + (L.n_raster + L.n_fill + L.n_quad) * 4ull, 256);
		L.items_domain = off;
		off = align_up(off + (L.n_dom_pre + L.n_dom_post + L.n_resolve) * 4ull, 256);
		L.items_up = off;
		off = align_up(off + (L.n_up + L.n_up_fill) * 4ull, 256);
		for (size_t i = 0; i < rounds[r].size(); i++) {
			const HostPass &hp = *rounds[r][i].pass;
			PassInputs &pi = pl[r][i];
			pi.words = off;
			off = align_up(off + hp.words.size() * 4, 256);
			pi.states = off;
			off = align_up(off + hp.states.size() * sizeof(DrawState), 256);
			pi.prims = off;
			off = align_up(off + hp.prims.size() * sizeof(PrimRecord), 256);
			pi.uploads = off;
			off = align_up(off + std::max<size_t>(hp.uploads.size(), 1) * sizeof(UploadInfo), 256);
		}
	}
	const size_t total_input = off;
	c.last_arena_bytes = total_input + max_scratch;

	if (input.fixed && total_input > input.size)
		return set_error(RDP_MI_ERROR_LIMIT, "the submission needs %zu bytes of packed inputs, more than the %zu the static "
		                 "array holds", total_input, input.size);
	if (scratch.fixed && max_scratch > scratch.size)
		return set_error(RDP_MI_ERROR_LIMIT, "the submission needs %zu bytes of scratch, more than the %zu the static array "
		                 "holds", max_scratch, scratch.size);
	int err;
	if ((err = staging.reserve(total_input))) return set_error(err, "staging allocation of %zu bytes failed", total_input);
	if ((err = input.reserve(total_input, c.stream)))
		return set_error(err, "input allocation of %zu bytes failed", total_input);
	if ((err = scratch.reserve(std::max<size_t>(max_scratch, 256), c.stream)))
		return set_error(err, "scratch allocation of %zu bytes failed", max_scratch);

	uint8_t *h = staging.ptr;
	uint8_t *d = input.ptr;
	launches.resize(rounds.size());
	// Per processor: the TMEM snapshots of its pass in the round before, for
	// an upscaled pass that follows it.
	std::map<const Processor *, uint16_t *> native_tmem;
	for (size_t r = 0; r < rounds.size(); r++) {
		RoundLayout &L = rl[r];
		PassDesc *descs = reinterpret_cast<PassDesc *>(h + L.descs);
		uint32_t *it_setup = reinterpret_cast<uint32_t *>(h + L.items_setup);
		uint32_t *it_tmem = reinterpret_cast<uint32_t *>(h + L.items_tmem);
		uint32_t *it_bin = reinterpret_cast<uint32_t *>(h + L.items_bin);
		uint32_t *it_raster = reinterpret_cast<uint32_t *>(h + L.items_raster);
		uint32_t *it_domain = reinterpret_cast<uint32_t *>(h + L.items_domain);
		uint32_t *it_up = reinterpret_cast<uint32_t *>(h + L.items_up);
		uint32_t ns = 0, nsp = 0, nt = 0, ntg = 0, nb = 0, ndp = 0, ndq = 0, nrs = 0;
		uint32_t upbase[kRasterClasses], nu[kRasterClasses];
		for (uint32_t c = 0, acc = 0; c < kRasterClasses; c++) {
			upbase[c] = acc;
			nu[c] = 0;
			acc += L.n_up_class[c];
		}
		std::map<const Processor *, uint16_t *> round_tmem;
		// Raster items: the 8x8-tile list by class, the fill items, then the
		// 16x16-tile list by class.
		uint32_t nr[kRasterClasses], rbase[kRasterClasses], nq[kRasterClasses], qbase[kRasterClasses];
		for (uint32_t c = 0, acc = 0, qacc = L.n_raster + L.n_fill; c < kRasterClasses; c++) {
			nr[c] = nq[c] = 0;
			rbase[c] = acc;
			acc += L.n_raster_class[c];
			qbase[c] = qacc;
			qacc += L.n_quad_class[c];
		}
		Launch &LA = launches[r];
		LA.descs = reinterpret_cast<const PassDesc *>(d + L.descs);
		LA.items[0] = reinterpret_cast<const uint32_t *>(d + L.items_setup);
		LA.items[1] = reinterpret_cast<const uint32_t *>(d + L.items_tmem);
		LA.tmem_scan_items = LA.items[1] + L.n_tmem_gather;
		LA.tmem_scan_n = L.n_tmem;
		LA.span_items = LA.items[0] + L.n_setup;
		LA.span_n = L.n_span;
		LA.items[2] = reinterpret_cast<const uint32_t *>(d + L.items_bin);
		LA.items[3] = reinterpret_cast<const uint32_t *>(d + L.items_raster);
		LA.n[0] = L.n_setup;
		LA.n[1] = L.n_tmem_gather;
		LA.n[2] = L.n_bin;
		LA.n[3] = L.n_raster;
		LA.fill_items = LA.items[3] + L.n_raster;
		LA.fill_n = L.n_fill;
		LA.prep_ahead = prep_ahead;
		LA.tmem_early = tmem_may_start_early(rounds, r);
		for (uint32_t c = 0; c < kRasterClasses; c++) {
			LA.raster_items[c] = LA.items[3] + rbase[c];
			LA.raster_n[c] = L.n_raster_class[c];
			LA.quad_items[c] = LA.items[3] + qbase[c];
			LA.quad_n[c] = L.n_quad_class[c];
			LA.up_items[c] = reinterpret_cast<const uint32_t *>(d + L.items_up) + upbase[c];
			LA.up_n[c] = L.n_up_class[c];
		}
		LA.up_fill_items = reinterpret_cast<const uint32_t *>(d + L.items_up) + L.n_up;
		LA.up_fill_n = L.n_up_fill;
		LA.domain_pre_items = reinterpret_cast<const uint32_t *>(d + L.items_domain);
		LA.domain_pre_n = L.n_dom_pre;
		LA.domain_post_items = LA.domain_pre_items + L.n_dom_pre;
		LA.domain_post_n = L.n_dom_post;
		LA.resolve_items = LA.domain_post_items + L.n_dom_post;
		LA.resolve_n = L.n_resolve;
		LA.raster_bytes = 0;
		LA.fill_bytes = 0;
		LA.fill_tiles = 0;
		for (uint32_t c = 0; c < kRasterClasses; c++) LA.class_bytes[c] = 0;
		LA.passes = (uint32_t)rounds[r].size();
		for (size_t i = 0; i < rounds[r].size(); i++) {
			const uint64_t fb_bytes = framebuffer_bytes(*rounds[r][i].pass);
			LA.raster_bytes += fb_bytes;
			if (pl[r][i].fill) {
				LA.fill_bytes += fb_bytes;
				if (!rounds[r][i].pass->prims.empty())
					LA.fill_tiles += ((rounds[r][i].pass->fb_width + 7) / 8) * ((rounds[r][i].pass->fb_height + 7) / 8);
			} else
				LA.class_bytes[std::min(rounds[r][i].pass->raster_class, kRasterClasses - 1)] += fb_bytes;
			const HostPass &hp = *rounds[r][i].pass;
			Processor &pr = *rounds[r][i].proc;
			PassInputs &pi = pl[r][i];
			memcpy(h + pi.words, hp.words.data(), hp.words.size() * 4);
			memcpy(h + pi.states, hp.states.data(), hp.states.size() * sizeof(DrawState));
			memcpy(h + pi.prims, hp.prims.data(), hp.prims.size() * sizeof(PrimRecord));
			if (!hp.uploads.empty())
				memcpy(h + pi.uploads, hp.uploads.data(), hp.uploads.size() * sizeof(UploadInfo));

			PassDesc &D = descs[i];
			memset(&D, 0, sizeof(D));
			D.words = reinterpret_cast<const uint32_t *>(d + pi.words);
			D.states = reinterpret_cast<const DrawState *>(d + pi.states);
			D.prims = reinterpret_cast<const PrimRecord *>(d + pi.prims);
			D.uploads = reinterpret_cast<const UploadInfo *>(d + pi.uploads);
			uint8_t *sb = scratch.ptr + pi.scratch_base;
			D.rprims = reinterpret_cast<RasterPrim *>(sb);
			D.spans = reinterpret_cast<SpanSetup *>(sb + pi.sl.setups);
			D.tilerows = reinterpret_cast<uint32_t *>(sb + pi.sl.setups + pi.sl.spans);
			D.bin_fine = reinterpret_cast<uint64_t *>(sb + pi.sl.setups + pi.sl.spans + pi.sl.tilerows);
			D.tmem_blocks = reinterpret_cast<uint32_t *>(sb + pi.sl.setups + pi.sl.spans + pi.sl.tilerows + pi.sl.bins +
			                                             pi.sl.tmem);
			D.tmem_instances = reinterpret_cast<uint16_t *>(sb + pi.sl.setups + pi.sl.spans + pi.sl.tilerows + pi.sl.bins);
			{
				uint8_t *tail = sb + pi.sl.setups + pi.sl.spans + pi.sl.tilerows + pi.sl.bins + pi.sl.tmem + pi.sl.tmem_blocks;
				D.psetups = reinterpret_cast<PrimSetup *>(tail);
				D.tilerow_prim = reinterpret_cast<uint16_t *>(tail + pi.sl.psetups);
			}
			D.rdram = pr.d_rdram;
			D.hidden = pr.d_hidden;
			D.tmem = pr.d_tmem;
			D.wmask = pr.coherent ? pr.d_wmask : nullptr;
			D.rdram_mask = pr.rdram_size - 1;
			D.num_prims = (uint32_t)hp.prims.size();
			D.num_uploads = (uint32_t)hp.uploads.size();
			D.fb_addr = hp.fb_addr;
			D.fb_width = hp.fb_width;
			D.fb_height = hp.fb_height;
			D.fb_fmt = hp.fb_fmt;
			D.depth_addr = hp.depth_addr;
			D.tiles_x = (hp.fb_width + 7) / 8;
			D.tiles_y = (hp.fb_height + 7) / 8;
			D.num_spans = hp.num_spans;
			D.num_tilerows = hp.num_tilerows;
			D.num_groups = bin_groups(hp.prims.size());
			D.scale_log2 = hp.scale_log2;
			D.up_rdram = pr.d_up_rdram;
			D.up_hidden = pr.d_up_hidden;
			D.up_ref = pr.d_up_ref;
			D.domain_height = hp.domain_height;
			D.domain_log2 = pr.up_log2;
			D.up_wmask = pr.d_up_wmask;
			D.domain_flags = (hp.ssaa_clear ? DOMAIN_SSAA_CLEAR : 0u) | (hp.ssaa_resolve ? DOMAIN_SSAA_RESOLVE : 0u) |
			                 (pr.ssaa_dither ? DOMAIN_SSAA_DITHER : 0u);
			D.resolve_rdram = pr.d_rdram;
			if (hp.scale_log2) {
				// The upscaled pass renders into the sample planes with the
				// TMEM snapshots of its native pass.
				auto it = native_tmem.find(&pr);
				if (it == native_tmem.end())
					return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "upscaled pass without its native pass");
				D.rdram = pr.d_up_rdram;
				D.hidden = pr.d_up_hidden;
				D.wmask = nullptr;
				D.tmem_instances = it->second;
			} else
				round_tmem[&pr] = D.tmem_instances;

			const uint32_t pidx = (uint32_t)i;
			for (uint32_t g = 0; g < (D.num_prims + 63) / 64; g++) it_setup[ns++] = (pidx << 7) | g;
			for (uint32_t b = 0; b < (D.num_tilerows + 7) / 8; b++) it_setup[L.n_setup + nsp++] = (pidx << 17) | b;
			if (!hp.scale_log2) {
				for (uint32_t k = 0; k < 8; k++) it_tmem[L.n_tmem_gather + nt++] = (pidx << 12) | k;
				for (uint32_t b = 0; b < (uint32_t)(hp.uploads.size() + 31) / 32; b++)
					for (uint32_t k = 0; k < 8; k++) it_tmem[ntg++] = (pidx << 12) | (b << 3) | k;
			}
			if (D.num_prims) {
				for (uint32_t y = 0; y < D.tiles_y; y++)
					for (uint32_t g = 0; g < D.num_groups / 8; g++) it_bin[nb++] = (pidx << 17) | (y << 4) | g;
			}
			if (hp.domain_pre)
				for (uint32_t b = 0; b < domain_blocks(hp); b++) it_domain[ndp++] = (pidx << 17) | b;
			if (hp.domain_post)
				for (uint32_t b = 0; b < domain_blocks(hp); b++) it_domain[L.n_dom_pre + ndq++] = (pidx << 17) | b;
			if (hp.ssaa_resolve)
				for (uint32_t b = 0; b < resolve_blocks(hp); b++)
					it_domain[L.n_dom_pre + L.n_dom_post + nrs++] = (pidx << 17) | b;
		}
		// Setup, span and bin items: the passes above class 0 first.
		{
			auto heavy = [&](uint32_t pidx) { return rounds[r][pidx].pass->raster_class > 0; };
			LA.setup_hi = (uint32_t)(std::stable_partition(it_setup, it_setup + L.n_setup,
			                                               [&](uint32_t it) { return heavy(it >> 7); }) - it_setup);
			uint32_t *sp = it_setup + L.n_setup;
			LA.span_hi = (uint32_t)(std::stable_partition(sp, sp + L.n_span, [&](uint32_t it) { return heavy(it >> 17); }) - sp);
			LA.bin_hi = (uint32_t)(std::stable_partition(it_bin, it_bin + L.n_bin, [&](uint32_t it) { return heavy(it >> 17); }) - it_bin);
		}
		// Raster items, per class, passes with the most primitives per tile
		// first: their tiles have the longest chains, so they start early
		// and the short tiles of light passes fill the end of the launch.
		std::vector<uint32_t> order(rounds[r].size());
		for (uint32_t i = 0; i < (uint32_t)order.size(); i++) order[i] = i;
		auto density = [&](uint32_t i) {
			const HostPass &hp = *rounds[r][i].pass;
			const uint64_t tiles = (uint64_t)((hp.fb_width + 7) / 8) * ((hp.fb_height + 7) / 8);
			return tiles ? (double)hp.prims.size() / (double)tiles : 0.0;
		};
		std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return density(a) > density(b); });
		uint32_t nf = 0, nuf = 0;
		for (uint32_t pidx : order) {
			const HostPass &hp = *rounds[r][pidx].pass;
			if (hp.prims.empty()) continue;
			const uint32_t tiles_x = (hp.fb_width + 7) / 8, tiles_y = (hp.fb_height + 7) / 8;
			if (hp.scale_log2 && pl[r][pidx].fill) {
				uint32_t *out = it_up + L.n_up;
				for (uint32_t y = 0; y < tiles_y; y++)
					for (uint32_t x = 0; x < (tiles_x + kFillTilesPerItem - 1) / kFillTilesPerItem; x++)
						out[nuf++] = (pidx << 20) | (x << 10) | y;
				continue;
			}
			if (hp.scale_log2) {
				const uint32_t cls = std::min(hp.raster_class, kRasterClasses - 1);
				uint32_t *out = it_up + upbase[cls];
				for (uint32_t y = 0; y < tiles_y; y++)
					for (uint32_t x = 0; x < tiles_x; x++) out[nu[cls]++] = (pidx << 20) | (x << 10) | y;
				continue;
			}
			if (pl[r][pidx].fill) {
				uint32_t *out = it_raster + L.n_raster;
				for (uint32_t y = 0; y < tiles_y; y++)
					for (uint32_t x = 0; x < (tiles_x + kFillTilesPerItem - 1) / kFillTilesPerItem; x++)
						out[nf++] = (pidx << 14) | (x << 7) | y;
				continue;
			}
			const uint32_t cls = std::min(hp.raster_class, kRasterClasses - 1);
			uint32_t *out = it_raster + rbase[cls];
			for (uint32_t y = 0; y < tiles_y; y++)
				for (uint32_t x = 0; x < tiles_x; x++) out[nr[cls]++] = (pidx << 14) | (x << 7) | y;
			uint32_t *qout = it_raster + qbase[cls];
			for (uint32_t y = 0; y < (tiles_y + 1) / 2; y++)
				for (uint32_t x = 0; x < (tiles_x + 1) / 2; x++) qout[nq[cls]++] = (pidx << 14) | (x << 7) | y;
		}
		native_tmem = std::move(round_tmem);
	}

	HIP_CHECK(hipMemcpyAsync(d, h, total_input, hipMemcpyHostToDevice, c.stream));
	HIP_CHECK(hipEventRecord(c.staging_done, c.stream));
	c.staging_busy = true;
	return 0;
}

// Submit rounds of passes through a lane's arenas.  The caller owns the lane
// (for the context's own lane: holds the context lock).
// Arenas a submission uses in place of its lane's own.
struct Arenas {
	DeviceBuffer *input, *scratch;
	PinnedBuffer *staging;
};

static int submit_rounds(Lane &c, const std::vector<std::vector<PassRef>> &rounds, const Arenas *arenas = nullptr)
{
	std::vector<Launch> launches;
	int err = arenas ? prepare_rounds(c, rounds, *arenas->input, *arenas->scratch, launches, arenas->staging)
	                 : prepare_rounds(c, rounds, c.input, c.scratch, launches);
	if (err) return err;
	return execute_launches(launches, c, ctx().cfg);
}

static void build_round_list(Processor *const *procs, unsigned count, bool batched,
                             std::vector<std::vector<PassRef>> &rounds)
{
	rounds.clear();
	for (unsigned i = 0; i < count; i++) {
		auto &passes = procs[i]->fe.completed();
		for (size_t k = 0; k < passes.size(); k++) {
			if (!batched) {
				rounds.push_back({ { procs[i], &passes[k] } });
				continue;
			}
			if (rounds.size() <= k) rounds.resize(k + 1);
			rounds[k].push_back({ procs[i], &passes[k] });
		}
	}
}

// ---- host-coherent RDRAM ----
constexpr uint32_t kCoherentPage = 4096;

// Page-aligned runs [lo, hi) inside [0, size) covering the byte ranges
// (which may wrap at the RDRAM size).
static void page_runs(const std::vector<ByteRange> &ranges, uint32_t size, std::vector<ByteRange> &runs)
{
	runs.clear();
	const uint32_t pages = size / kCoherentPage;
	std::vector<uint8_t> mark(pages, 0);
	for (const ByteRange &r : ranges) {
		if (r.hi <= r.lo) continue;
		if (r.hi - r.lo >= size) {
			std::fill(mark.begin(), mark.end(), 1);
			break;
		}
		const uint64_t first = r.lo / kCoherentPage, last = (r.hi - 1) / kCoherentPage;
		for (uint64_t pg = first; pg <= last; pg++) mark[pg % pages] = 1;
	}
	for (uint32_t pg = 0; pg < pages;) {
		if (!mark[pg]) {
			pg++;
			continue;
		}
		uint32_t e = pg;
		while (e < pages && mark[e]) e++;
		runs.push_back({ (uint64_t)pg * kCoherentPage, (uint64_t)e * kCoherentPage });
		pg = e;
	}
}

// Merge the write-back in flight into the caller's memory: bytes the raster
// marked take the device value.
static int coherent_complete(Processor &p)
{
	std::lock_guard<std::mutex> g(p.wb_lock);
	if (!p.wb_pending) return 0;
	HIP_CHECK(hipEventSynchronize(p.wb_event));
	uint8_t *host = p.user_rdram;
	for (const ByteRange &r : p.wb_runs) {
		const uint64_t *d = reinterpret_cast<const uint64_t *>(p.wb_data.ptr + r.lo);
		const uint64_t *m = reinterpret_cast<const uint64_t *>(p.wb_mask.ptr + r.lo);
		uint64_t *h = reinterpret_cast<uint64_t *>(host + r.lo);
		for (uint64_t i = 0; i < (r.hi - r.lo) / 8; i++) {
			uint64_t hv;
			memcpy(&hv, h + i, 8);  // the caller's pointer may be unaligned
			hv = (hv & ~m[i]) | (d[i] & m[i]);
			memcpy(h + i, &hv, 8);
		}
	}
	p.wb_pending = false;
	return 0;
}

// Copy in the pages the passes read, before their launches.
static int coherent_pull(Processor &p, const std::vector<HostPass> &passes, hipStream_t st)
{
	if (int err = coherent_complete(p)) return err;
	std::vector<ByteRange> ranges, runs;
	for (const HostPass &hp : passes) {
		pass_write_ranges(hp, ranges);
		for (const UploadInfo &u : hp.uploads) {
			ByteRange b;
			if (!upload_read_range(u, b)) b = { 0, p.rdram_size };
			ranges.push_back(b);
		}
	}
	page_runs(ranges, p.rdram_size, runs);
	for (const ByteRange &r : runs)
		HIP_CHECK(hipMemcpyAsync(p.d_rdram + r.lo, p.user_rdram + r.lo, r.hi - r.lo, hipMemcpyHostToDevice, st));
	return 0;
}

// After the launches: copy out the written pages and their mask, clear the
// mask, and leave the merge to coherent_complete.
static int coherent_push(Processor &p, const std::vector<HostPass> &passes, hipStream_t st)
{
	std::vector<ByteRange> ranges;
	for (const HostPass &hp : passes) pass_write_ranges(hp, ranges);
	std::lock_guard<std::mutex> g(p.wb_lock);
	page_runs(ranges, p.rdram_size, p.wb_runs);
	if (p.wb_runs.empty()) return 0;
	for (const ByteRange &r : p.wb_runs) {
		const size_t n = r.hi - r.lo;
		HIP_CHECK(hipMemcpyAsync(p.wb_data.ptr + r.lo, p.d_rdram + r.lo, n, hipMemcpyDeviceToHost, st));
		HIP_CHECK(hipMemcpyAsync(p.wb_mask.ptr + r.lo, p.d_wmask + r.lo, n, hipMemcpyDeviceToHost, st));
		HIP_CHECK(hipMemsetAsync(p.d_wmask + r.lo, 0, n, st));
	}
	HIP_CHECK(hipEventRecord(p.wb_event, st));
	p.wb_pending = true;
	return 0;
}

// Submit every completed pass of the given processors (plain path: one
// processor, one pass per round) on a lane (default: the context's).
static int submit_processors(Processor *const *procs, unsigned count, bool batched, Lane *lane = nullptr,
                             const Arenas *arenas = nullptr)
{
	Lane &c = lane ? *lane : ctx();
	std::vector<std::vector<PassRef>> rounds;
	for (unsigned i = 0; i < count; i++) {
		auto &passes = procs[i]->fe.completed();
		for (size_t k = 0; k < passes.size(); k++) {
			if (!batched) {
				rounds.push_back({ { procs[i], &passes[k] } });
				continue;
			}
			if (rounds.size() <= k) rounds.resize(k + 1);
			rounds[k].push_back({ procs[i], &passes[k] });
		}
	}
	int ret = 0;
	for (unsigned i = 0; i < count && !ret; i++)
		if (procs[i]->coherent) ret = coherent_pull(*procs[i], procs[i]->fe.completed(), c.stream);
	if (ret) {
	} else if (batched) {
		// Keep each launch's pass count within the work-item encoding.
		ret = submit_rounds(c, rounds, arenas);
	} else {
		// One pass per submission, except that an upscaled pass goes with its
		// native pass: it reads the TMEM snapshots in that pass's scratch.
		for (size_t k = 0; k < rounds.size() && !ret;) {
			std::vector<std::vector<PassRef>> group{ rounds[k++] };
			while (k < rounds.size() && rounds[k][0].pass->scale_log2) group.push_back(rounds[k++]);
			ret = submit_rounds(c, group, arenas);
		}
	}
	// A batched submission that does not fit fixed arenas fails before its
	// first copy or launch: its passes stay pending for another entry point.
	if (ret == RDP_MI_ERROR_LIMIT && batched && arenas) return ret;
	for (unsigned i = 0; i < count && !ret; i++)
		if (procs[i]->coherent) ret = coherent_push(*procs[i], procs[i]->fe.completed(), c.stream);
	for (unsigned i = 0; i < count; i++) procs[i]->fe.completed().clear();
	return ret;
}

// Stream-ordered synchronous copy (the context stream does not synchronise
// with the null stream).
static int copy_sync(void *dst, const void *src, size_t size, hipMemcpyKind kind)
{
	Context &c = ctx();
	std::lock_guard<std::mutex> g(c.lock);
	HIP_CHECK(hipMemcpyAsync(dst, src, size, kind, c.stream));
	HIP_CHECK(hipStreamSynchronize(c.stream));
	return 0;
}

// ---------------------------------------------------------------------------
// Timeline completion (the reference's FenceExecutor on a WorkerThread,
// rdp_device.cpp:1184-1263, worker_thread.hpp:39-126).  A signal records an
// event on the context stream behind the processor's work and queues it here;
// one thread waits for the events in order and publishes each value.
// ---------------------------------------------------------------------------
static void publish_timeline(Processor &p, uint64_t value, int err)
{
	std::lock_guard<std::mutex> g(p.tl_lock);
	if (err && !p.timeline_error) p.timeline_error = err;
	if (value > p.timeline_done) p.timeline_done = value;
	p.tl_cv.notify_all();
}

class TimelineWorker {
public:
	struct Item {
		Processor *proc;
		uint64_t value;
		hipEvent_t ev;
		int err;
	};

	// An event from the pool, or a new one (null if creation fails).
	hipEvent_t acquire()
	{
		{
			std::lock_guard<std::mutex> g(m);
			if (!pool.empty()) {
				hipEvent_t ev = pool.back();
				pool.pop_back();
				return ev;
			}
		}
		hipEvent_t ev = nullptr;
		if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) return nullptr;
		return ev;
	}
	void release(hipEvent_t ev)
	{
		std::lock_guard<std::mutex> g(m);
		pool.push_back(ev);
	}
	void push(const Item &it)
	{
		std::lock_guard<std::mutex> g(m);
		if (!started) {
			started = true;
			std::thread([this] { loop(); }).detach();
		}
		queue.push_back(it);
		cv.notify_all();
	}

private:
	std::mutex m;
	std::condition_variable cv;
	std::deque<Item> queue;
	std::vector<hipEvent_t> pool;
	bool started = false;

	void loop()
	{
		for (;;) {
			Item it;
			{
				std::unique_lock<std::mutex> g(m);
				cv.wait(g, [this] { return !queue.empty(); });
				it = queue.front();
				queue.pop_front();
			}
			const hipError_t e = hipEventSynchronize(it.ev);
			release(it.ev);
			publish_timeline(*it.proc, it.value, it.err ? it.err : (e != hipSuccess ? RDP_MI_ERROR_DEVICE : 0));
		}
	}
};

static TimelineWorker &timeline_worker()
{
	// Never destroyed: its thread may still be parked on the queue at exit.
	static TimelineWorker *w = new TimelineWorker();
	return *w;
}

// Hand the processor's completed passes to the device (caller holds p.lock).
static int submit_pending(Processor &p)
{
	if (p.fe.completed().empty()) return 0;
	Context &c = ctx();
	std::lock_guard<std::mutex> g(c.lock);
	Processor *pp = &p;
	const int err = submit_processors(&pp, 1, false);
	// A failed submission from the ring thread (idle or pass-count submits)
	// is reported by the next timeline signal, with its message.
	if (err && !p.submit_error) {
		p.submit_error = err;
		p.submit_error_msg = g_last_error;
	}
	return err;
}

// MetaFlush: end the open pass and submit (caller holds p.lock).
static int flush_now(Processor &p)
{
	p.fe.flush();
	return submit_pending(p);
}

// MetaSignalTimeline (rdp_device.cpp:926-934): flush, then publish `value`
// once the device has finished everything submitted so far.  Without an event
// the stream is waited for here, so a value is never published early.
static void signal_now(Processor &p, uint64_t value)
{
	const int err = flush_now(p);
	if (p.submit_error) {
		std::lock_guard<std::mutex> g(p.tl_lock);
		if (!p.timeline_error) {
			p.timeline_error = p.submit_error;
			p.timeline_error_msg = p.submit_error_msg;
		}
	}
	Context &c = ctx();
	TimelineWorker &w = timeline_worker();
	hipEvent_t ev = w.acquire();
	if (ev) {
		std::lock_guard<std::mutex> g(c.lock);
		if (hipEventRecord(ev, c.stream) != hipSuccess) {
			w.release(ev);
			ev = nullptr;
		}
	}
	if (ev) {
		w.push({ &p, value, ev, err });
		return;
	}
	hipError_t e;
	{
		std::lock_guard<std::mutex> g(c.lock);
		e = hipStreamSynchronize(c.stream);
	}
	publish_timeline(p, value, err ? err : (e != hipSuccess ? RDP_MI_ERROR_DEVICE : 0));
}

// Number of words of a command in the ring, meta commands included.
static unsigned ring_command_length(uint32_t w0)
{
	const uint32_t op = (w0 >> 24) & 63;
	if (op == kMetaSignal) return 3;
	if (op == kMetaFlush) return 1;
	return Frontend::command_length(w0);
}

// Decode complete commands in order (caller holds p.lock).  Completed passes
// are handed to the device once enough queue up, bounding host memory.
static void run_commands(Processor &p, const uint32_t *w, size_t n)
{
	size_t i = 0;
	while (i < n) {
		const uint32_t op = (w[i] >> 24) & 63;
		const unsigned len = ring_command_length(w[i]);
		if (len == 0) {
			i++;
			continue;
		}
		if (i + len > n) break;
		if (op == kMetaSignal)
			signal_now(p, (uint64_t)w[i + 1] | ((uint64_t)w[i + 2] << 32));
		else if (op == kMetaFlush)
			(void)flush_now(p);
		else {
			p.fe.command(w + i);
			if (p.fe.completed().size() >= 32 && !p.hold.load()) (void)submit_pending(p);
		}
		i += len;
	}
}

// The ring thread (CommandRing::thread_loop, command_ring.cpp:86-134).  When
// no command arrives for 500 us while passes are waiting, it hands them to
// the device (the reference's MetaIdle).
static void ring_loop(Processor *p)
{
	CommandRing &r = *p->ring;
	std::vector<uint32_t> batch;
	for (;;) {
		bool idle = false;
		{
			std::unique_lock<std::mutex> g(r.lock);
			auto has_work = [&r] { return r.write_count > r.read_count || r.stop; };
			if (r.pending)
				idle = !r.cv.wait_for(g, std::chrono::microseconds(500), has_work);
			else
				r.cv.wait(g, has_work);
			if (!idle) {
				if (r.write_count == r.read_count) break;  // stop requested, ring empty
				const size_t n = (size_t)(r.write_count - r.read_count);
				batch.resize(n);
				for (size_t k = 0; k < n; k++) batch[k] = r.words[(r.read_count + k) & (kRingWords - 1)];
				r.read_count += n;
				r.cv.notify_all();  // space for producers
			}
		}
		bool pending;
		{
			std::lock_guard<std::mutex> g(p->lock);
			if (idle) {
				if (!p->hold.load()) (void)submit_pending(*p);
			} else
				run_commands(*p, batch.data(), batch.size());
			pending = !p->fe.completed().empty() && !p->hold.load();
		}
		std::lock_guard<std::mutex> g(r.lock);
		r.pending = pending;
		if (!idle) {
			r.completed_count = r.read_count;
			r.cv.notify_all();  // drains
		}
	}
}

static CommandRing &ensure_ring(Processor &p)
{
	std::lock_guard<std::mutex> g(p.ring_create);
	if (!p.ring) {
		p.ring.reset(new CommandRing());
		p.ring->thr = std::thread(ring_loop, &p);
	}
	return *p.ring;
}

// Copy complete commands (n <= kRingWords words) into the ring.
static void ring_push(Processor &p, const uint32_t *w, size_t n)
{
	CommandRing &r = ensure_ring(p);
	std::unique_lock<std::mutex> g(r.lock);
	r.cv.wait(g, [&] { return r.write_count + n <= r.read_count + kRingWords; });
	for (size_t k = 0; k < n; k++) r.words[(r.write_count + k) & (kRingWords - 1)] = w[k];
	r.write_count += n;
	r.cv.notify_all();
}

// Wait until the ring thread has decoded everything enqueued so far
// (CommandRing::drain, command_ring.cpp:63-69).
static void ring_drain(Processor &p)
{
	CommandRing *r;
	{
		std::lock_guard<std::mutex> g(p.ring_create);
		r = p.ring.get();
	}
	if (!r) return;
	std::unique_lock<std::mutex> g(r->lock);
	r->cv.wait(g, [r] { return r->completed_count == r->write_count; });
}

static bool has_ring(Processor &p)
{
	std::lock_guard<std::mutex> g(p.ring_create);
	return p.ring != nullptr;
}

// Send a meta command in-band when the processor has a ring, else run it
// on the calling thread.
static void meta_command(Processor &p, const uint32_t *w, size_t n)
{
	if (has_ring(p)) {
		ring_push(p, w, n);
		return;
	}
	std::lock_guard<std::mutex> g(p.lock);
	run_commands(p, w, n);
}

static void stop_ring(Processor &p)
{
	CommandRing *r;
	{
		std::lock_guard<std::mutex> g(p.ring_create);
		r = p.ring.get();
	}
	if (!r) return;
	{
		std::lock_guard<std::mutex> g(r->lock);
		r->stop = true;
		r->cv.notify_all();
	}
	if (r->thr.joinable()) r->thr.join();
}

// A recorded submission: device-resident inputs and scratch for a fixed set
// of passes, replayable as direct launches or as one captured hipGraph.
struct Recording {
	std::vector<Launch> launches;
	DeviceBuffer input, scratch;
	LaunchConfig cfg;
	std::vector<std::pair<LaunchConfig, float>> tuned;  // every autotune candidate and its time
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	uint64_t raster_bytes = 0;
	uint32_t passes = 0;
	// The recorded processors' device state, for trial runs (autotune,
	// profile) that must leave it as they found it.
	std::vector<Processor *> procs;
	DeviceBuffer saved;

	size_t state_bytes() const
	{
		size_t n = 0;
		for (const Processor *p : procs) n += (size_t)p->rdram_size + p->rdram_size / 2 + 4096;
		return n;
	}
	// Copies every processor's RDRAM, hidden RDRAM and TMEM aside (restore
	// = false) or back (restore = true), in stream order.
	int copy_state(hipStream_t st, bool restore)
	{
		if (!restore && saved.reserve(state_bytes(), st)) return set_error(RDP_MI_ERROR_OUT_OF_MEMORY, "state save buffer");
		size_t off = 0;
		for (Processor *p : procs) {
			void *dev[3] = { p->d_rdram, p->d_hidden, p->d_tmem };
			const size_t sz[3] = { p->rdram_size, p->rdram_size / 2, 4096 };
			for (int k = 0; k < 3; k++) {
				void *dst = restore ? dev[k] : saved.ptr + off, *src = restore ? saved.ptr + off : dev[k];
				HIP_CHECK(hipMemcpyAsync(dst, src, sz[k], hipMemcpyDeviceToDevice, st));
				off += sz[k];
			}
		}
		return 0;
	}
};

// ---------------------------------------------------------------------------
// Serving: continuous batching of frame requests from many processors.
//
// Requests (one frame of commands for one processor, with an arrival time)
// queue per processor.  Three kinds of threads move them along:
//   - the scheduler admits every arrived request whose processor has nothing
//     in flight, in round-robin order, while the estimated device bytes of
//     admitted-but-unsubmitted frames (packed inputs plus kernel scratch) fit
//     in the HBM budget, and hands each admitted frame to the decode pool at
//     once;
//   - decode workers run the host front end of one frame each (front ends
//     are per processor, so frames decode in parallel) and mark it ready;
//   - each submission lane's submitter thread takes every ready frame (up
//     to max_batch, within the budget) whenever its lane is free, and
//     submits their passes as one batched launch set on that lane, so a
//     batch can start on one lane while a long one still runs on another
//     (launching kernels costs host time, so lanes submit in parallel);
//   - one completion thread per lane waits on its batches' events in order,
//     stamps completion times and frees the processors and the lane.
// Results are identical to rendering each frame alone: a processor has at
// most one frame in flight and processors share launches, never state.
// ---------------------------------------------------------------------------
class ThreadPool {
public:
	explicit ThreadPool(unsigned n)
	{
		for (unsigned i = 0; i < n; i++) workers.emplace_back([this] { run(); });
	}
	~ThreadPool()
	{
		{
			std::lock_guard<std::mutex> g(m);
			stop = true;
		}
		cv.notify_all();
		for (auto &t : workers) t.join();
	}
	void post(std::function<void()> job)
	{
		{
			std::lock_guard<std::mutex> g(m);
			queue.push_back(std::move(job));
		}
		cv.notify_one();
	}

private:
	void run()
	{
		for (;;) {
			std::function<void()> job;
			{
				std::unique_lock<std::mutex> g(m);
				cv.wait(g, [this] { return stop || !queue.empty(); });
				if (queue.empty()) return;
				job = std::move(queue.front());
				queue.pop_front();
			}
			job();
		}
	}
	std::vector<std::thread> workers;
	std::deque<std::function<void()>> queue;
	std::mutex m;
	std::condition_variable cv;
	bool stop = false;
};

struct ServeRequest {
	uint64_t ticket;
	uint32_t instance;
	double arrival_us;
	std::vector<uint32_t> words;
};

struct ServeResult {
	uint64_t ticket;
	uint32_t instance;
	double arrival_us, admit_us, submit_us, complete_us;
	uint64_t digest;  // state digest after the frame (when enabled)
	int32_t status;   // 0, or the error of the frame's batch
};

struct Server {
	std::vector<Processor *> procs;
	size_t budget = 0;
	unsigned max_batch = 0;
	unsigned decode_threads = 0;
	std::unique_ptr<ThreadPool> pool;

	std::mutex m;
	std::condition_variable cv;  // scheduler: arrivals, freed processors, freed budget
	std::condition_variable submit_cv;  // submitter: ready frames, free lanes
	std::vector<std::deque<ServeRequest>> queues;
	std::vector<bool> busy;
	uint64_t next_ticket = 1;
	size_t outstanding = 0;    // submitted by the caller, not completed
	size_t pending_bytes = 0;  // admitted, not yet submitted to the GPU
	size_t decoding = 0;       // admitted, not yet decoded
	double batch_window_us = 0;  // how long ready frames may wait for more (0: submit at once)
	std::vector<ServeResult> results;
	bool started = false, stopping = false;
	int error = 0;
	std::string error_msg;
	std::chrono::steady_clock::time_point t0;

	struct Ready {
		ServeResult r;
		size_t bytes;
		double ready_us;
	};
	std::deque<Ready> ready;

	struct Batch {
		hipEvent_t ev = nullptr;
		std::vector<ServeResult> reqs;
	};
	struct LaneState {
		Lane lane;
		hipEvent_t done_ev = nullptr;  // completion of the lane's batch in flight
		bool busy = false;
		// Per-frame digests: targets staged in pinned memory, results copied
		// back behind the batch (one batch in flight per lane).
		DigestTarget *h_targets = nullptr, *d_targets = nullptr;
		uint64_t *h_digests = nullptr, *d_digests = nullptr;
		std::deque<Batch> inflight;
		std::condition_variable cv;
		std::thread completer, submitter;
	};
	std::vector<std::unique_ptr<LaneState>> lanes;
	std::thread scheduler;
	unsigned rr = 0;
	bool digests = false;
	// Statistics: batches, frames, largest batch, largest estimated and
	// actual device bytes of one batch.
	uint64_t n_batches = 0, n_frames = 0, max_frames = 0, max_est_bytes = 0, max_arena_bytes = 0;
	double submit_time_us = 0, max_submit_us = 0;  // host time inside batched submissions

	double now_us() const { return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count(); }

	// Device bytes a frame's submission needs: its inputs (words, states,
	// primitive records) and scratch (setup records, spans, bins, TMEM
	// snapshots), estimated from the command words before decoding.
	static size_t estimate_bytes(const ServeRequest &r) { return r.words.size() * 4 * 24 + (1u << 20); }

	bool done() const { return stopping && outstanding == 0; }
	// With a batching window, a free lane takes the ready frames only once
	// max_batch are ready, nothing admitted is still decoding, or the oldest
	// has waited the window; returns the microseconds left otherwise.
	double window_left(double now) const
	{
		if (batch_window_us <= 0 || ready.size() >= max_batch || decoding == 0) return 0;
		return std::max(0.0, ready.front().ready_us + batch_window_us - now);
	}
	int queue_digests(LaneState &ls, const std::vector<Processor *> &ps);
	void decode(ServeRequest req, size_t bytes, double admit);
	void scheduler_loop();
	void submitter_loop(LaneState *ls);
	void completion_loop(LaneState &ls);
};

static DigestTarget digest_target(const Processor &p)
{
	return DigestTarget{ reinterpret_cast<const uint32_t *>(p.d_rdram), p.d_hidden, p.d_tmem, p.rdram_size / 4,
		                 p.rdram_size / 2 };
}

static uint32_t digest_words(const Processor &p) { return p.rdram_size / 4 + p.rdram_size / 8 + 1024; }

// Queue the digest of every processor of the lane's batch behind its
// launches, and the copy of the digests to the lane's pinned results.
int Server::queue_digests(LaneState &ls, const std::vector<Processor *> &ps)
{
	uint32_t max_words = 0;
	for (size_t k = 0; k < ps.size(); k++) {
		ls.h_targets[k] = digest_target(*ps[k]);
		max_words = std::max(max_words, digest_words(*ps[k]));
	}
	hipStream_t st = ls.lane.stream;
	HIP_CHECK(hipMemcpyAsync(ls.d_targets, ls.h_targets, ps.size() * sizeof(DigestTarget), hipMemcpyHostToDevice, st));
	HIP_CHECK(launch_digest(ls.d_targets, (uint32_t)ps.size(), max_words, ls.d_digests, st));
	HIP_CHECK(hipMemcpyAsync(ls.h_digests, ls.d_digests, ps.size() * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
	return 0;
}

void Server::decode(ServeRequest req, size_t bytes, double admit)
{
	Processor *p = procs[req.instance];
	const std::vector<uint32_t> &w = req.words;
	{
		std::lock_guard<std::mutex> pg(p->lock);
		size_t i = 0;
		while (i < w.size()) {
			const unsigned len = Frontend::command_length(w[i]);
			if (len == 0) {
				i++;
				continue;
			}
			if (i + len > w.size()) break;
			p->fe.command(w.data() + i);
			i += len;
		}
		p->fe.flush();
	}
	std::lock_guard<std::mutex> g(m);
	ready.push_back(Ready{ ServeResult{ req.ticket, req.instance, req.arrival_us, admit, 0.0, 0.0, 0, 0 }, bytes,
	                       now_us() });
	decoding--;
	submit_cv.notify_all();  // any lane's submitter may be the free one
}

void Server::scheduler_loop()
{
	std::unique_lock<std::mutex> g(m);
	for (;;) {
		if (done()) break;
		const double now = now_us();
		double next_arrival = 1e300;
		const unsigned n = (unsigned)procs.size();
		for (unsigned k = 0; k < n; k++) {
			const unsigned i = (rr + k) % n;
			if (busy[i] || queues[i].empty()) continue;
			const ServeRequest &r = queues[i].front();
			if (r.arrival_us > now) {
				next_arrival = std::min(next_arrival, r.arrival_us);
				continue;
			}
			const size_t est = estimate_bytes(r);
			if (pending_bytes != 0 && pending_bytes + est > budget) break;
			pending_bytes += est;
			decoding++;
			busy[i] = true;
			ServeRequest req = std::move(queues[i].front());
			queues[i].pop_front();
			const double admit = now_us();
			pool->post([this, req = std::move(req), est, admit]() mutable { decode(std::move(req), est, admit); });
		}
		rr = (rr + 1) % n;
		if (next_arrival < 1e300)
			cv.wait_for(g, std::chrono::microseconds((int64_t)std::max(1.0, next_arrival - now_us())));
		else
			cv.wait(g);
	}
}

// One submitter per lane: kernel launches cost host time, so lanes submit in
// parallel.
void Server::submitter_loop(LaneState *ls)
{
	std::unique_lock<std::mutex> g(m);
	for (;;) {
		submit_cv.wait(g, [&] { return done() || (!ready.empty() && !ls->busy); });
		if (ready.empty() && done()) break;
		if (ready.empty() || ls->busy) continue;
		if (const double left = window_left(now_us())) {
			submit_cv.wait_for(g, std::chrono::microseconds((int64_t)std::max(1.0, left)));
			continue;
		}

		// Take ready frames in order, within max_batch and the budget.
		std::vector<Ready> batch;
		size_t bytes = 0;
		while (!ready.empty() && batch.size() < max_batch) {
			if (!batch.empty() && bytes + ready.front().bytes > budget) break;
			bytes += ready.front().bytes;
			batch.push_back(ready.front());
			ready.pop_front();
		}
		ls->busy = true;
		if (!ready.empty()) submit_cv.notify_all();  // frames left for the other lanes
		g.unlock();

		std::vector<Processor *> ps;
		ps.reserve(batch.size());
		for (auto &r : batch) ps.push_back(procs[r.r.instance]);
		Batch b;
		const double t_sub = now_us();
		int err = submit_processors(ps.data(), (unsigned)ps.size(), true, &ls->lane);
		if (!err && digests) err = queue_digests(*ls, ps);
		// A lane holds one batch at a time, so its completion event is reused.
		b.ev = ls->done_ev;
		if (hipEventRecord(b.ev, ls->lane.stream) != hipSuccess) {
			if (!err) err = set_error(RDP_MI_ERROR_DEVICE, "event record failed");
			b.ev = nullptr;
		}
		if (err) {
			// Launches of the failed batch may be queued on any of the lane's
			// streams: wait for all of them before its processors are freed,
			// so no later frame of theirs overlaps them.
			for (hipStream_t st : { ls->lane.stream, ls->lane.side, ls->lane.prep })
				if (st) (void)hipStreamSynchronize(st);
			b.ev = nullptr;
		}
		const double submitted = now_us();
		for (auto &r : batch) {
			r.r.submit_us = submitted;
			r.r.status = err;
			b.reqs.push_back(r.r);
		}

		g.lock();
		if (err && !error) {
			error = err;
			error_msg = rdp_mi_last_error();
		}
		n_batches++;
		submit_time_us += submitted - t_sub;
		max_submit_us = std::max(max_submit_us, submitted - t_sub);
		n_frames += batch.size();
		max_frames = std::max<uint64_t>(max_frames, batch.size());
		max_est_bytes = std::max<uint64_t>(max_est_bytes, bytes);
		max_arena_bytes = std::max<uint64_t>(max_arena_bytes, ls->lane.last_arena_bytes);
		pending_bytes -= bytes;
		ls->inflight.push_back(std::move(b));
		ls->cv.notify_all();
		cv.notify_all();  // budget freed
	}
}

void Server::completion_loop(LaneState &ls)
{
	std::unique_lock<std::mutex> g(m);
	for (;;) {
		ls.cv.wait(g, [&] { return !ls.inflight.empty() || done(); });
		if (ls.inflight.empty()) break;
		hipEvent_t ev = ls.inflight.front().ev;
		g.unlock();
		if (ev) (void)hipEventSynchronize(ev);
		const double t = now_us();
		g.lock();
		Batch b = std::move(ls.inflight.front());
		ls.inflight.pop_front();
		for (size_t k = 0; k < b.reqs.size(); k++) {
			ServeResult &r = b.reqs[k];
			if (digests && r.status == 0) r.digest = ls.h_digests[k];
			r.complete_us = t;
			results.push_back(r);
			busy[r.instance] = false;
			outstanding--;
		}
		ls.busy = !ls.inflight.empty();
		cv.notify_all();
		submit_cv.notify_all();
		if (done()) {
			// Wake the other threads so they can see the end.
			for (auto &l : lanes) l->cv.notify_all();
		}
	}
}

}  // namespace rdp_mi

using namespace rdp_mi;

struct rdp_mi_server {
	Server s;
};

struct rdp_mi_recording {
	Recording r;
};

struct rdp_mi_processor {
	Processor p;
	explicit rdp_mi_processor(uint32_t size) : p(size) {}
};

// VIScanoutBuffer (rdp_device.hpp:53-58): pinned host pixels and a fence.
static constexpr size_t kScanoutBufferWords = 640 * 576;
struct rdp_mi_scanout_buffer {
	uint32_t *pixels = nullptr;
	size_t words = 0;  // capacity; grown for an upscaling processor's larger image
	hipEvent_t fence = nullptr;
	unsigned width = 0, height = 0;
	bool in_flight = false;
};

extern "C" {

const char *rdp_mi_last_error(void) { return g_last_error.c_str(); }

int rdp_mi_create(rdp_mi_processor **out, void *rdram_ptr, size_t rdram_offset, size_t rdram_size,
                  size_t hidden_rdram_size, uint32_t flags)
{
	if (!out) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "out is null");
	*out = nullptr;
	if (rdram_size < 0x10000 || (rdram_size & (rdram_size - 1)) || rdram_size > (1u << 30))
		return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "rdram_size must be a power of two");
	if (hidden_rdram_size != rdram_size / 2)
		return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "hidden_rdram_size must be rdram_size / 2");
	// The largest factor whose flag is set (rdp_device.cpp:188-195).
	const uint32_t up_log2 = (flags & RDP_MI_FLAG_UPSCALING_8X) ? 3 : (flags & RDP_MI_FLAG_UPSCALING_4X) ? 2
	                         : (flags & RDP_MI_FLAG_UPSCALING_2X) ? 1 : 0;
	if (up_log2 && rdram_ptr)
		return set_error(RDP_MI_ERROR_UNSUPPORTED, "upscaling with host-coherent RDRAM is not supported");
	const bool ssaa = (flags & RDP_MI_FLAG_SUPER_SAMPLED_READ_BACK) != 0;
	if (ssaa && !up_log2)
		return set_error(RDP_MI_ERROR_UNSUPPORTED, "super-sampled readback needs an upscaling flag");
	Context &c = ctx();
	std::lock_guard<std::mutex> g(c.lock);
	int err = c.init();
	if (err) return err;

	rdp_mi_processor *h = new (std::nothrow) rdp_mi_processor((uint32_t)rdram_size);
	if (!h) return set_error(RDP_MI_ERROR_OUT_OF_MEMORY, "out of host memory");
	Processor &p = h->p;
	auto fail = [&](int code, const char *msg) {
		if (p.d_up_rdram) (void)hipFree(p.d_up_rdram);
		if (p.d_up_hidden) (void)hipFree(p.d_up_hidden);
		if (p.d_up_ref) (void)hipFree(p.d_up_ref);
		if (p.d_up_wmask) (void)hipFree(p.d_up_wmask);
		if (p.d_rdram) (void)hipFree(p.d_rdram);
		if (p.d_hidden) (void)hipFree(p.d_hidden);
		if (p.d_tmem) (void)hipFree(p.d_tmem);
		if (p.d_wmask) (void)hipFree(p.d_wmask);
		if (p.wb_data.ptr) (void)hipHostFree(p.wb_data.ptr);
		if (p.wb_mask.ptr) (void)hipHostFree(p.wb_mask.ptr);
		if (p.wb_event) (void)hipEventDestroy(p.wb_event);
		delete h;
		return set_error(code, "%s", msg);
	};
	if (hipMalloc(&p.d_rdram, rdram_size) != hipSuccess) return fail(RDP_MI_ERROR_OUT_OF_MEMORY, "RDRAM allocation failed");
	if (hipMalloc(&p.d_hidden, hidden_rdram_size) != hipSuccess)
		return fail(RDP_MI_ERROR_OUT_OF_MEMORY, "hidden RDRAM allocation failed");
	if (hipMalloc(&p.d_tmem, 4096) != hipSuccess) return fail(RDP_MI_ERROR_OUT_OF_MEMORY, "TMEM allocation failed");
	if (rdram_ptr) {
		p.user_rdram = static_cast<uint8_t *>(rdram_ptr) + rdram_offset;
		p.coherent = true;
		if (hipMalloc(&p.d_wmask, rdram_size) != hipSuccess)
			return fail(RDP_MI_ERROR_OUT_OF_MEMORY, "RDRAM write mask allocation failed");
		if (hipMemsetAsync(p.d_wmask, 0, rdram_size, c.stream) != hipSuccess)
			return fail(RDP_MI_ERROR_DEVICE, "write mask clear failed");
		if (p.wb_data.reserve(rdram_size) || p.wb_mask.reserve(rdram_size))
			return fail(RDP_MI_ERROR_OUT_OF_MEMORY, "write-back buffers failed");
		if (hipEventCreateWithFlags(&p.wb_event, hipEventDisableTiming) != hipSuccess)
			return fail(RDP_MI_ERROR_DEVICE, "write-back event creation failed");
		if (hipMemcpyAsync(p.d_rdram, p.user_rdram, rdram_size, hipMemcpyHostToDevice, c.stream) != hipSuccess)
			return fail(RDP_MI_ERROR_DEVICE, "RDRAM upload failed");
	} else {
		p.rdram_mirror.assign(rdram_size, 0);
		if (hipMemsetAsync(p.d_rdram, 0, rdram_size, c.stream) != hipSuccess)
			return fail(RDP_MI_ERROR_DEVICE, "RDRAM clear failed");
	}
	p.hidden_mirror.assign(hidden_rdram_size, 0x03);
	if (hipMemsetAsync(p.d_hidden, 0x03, hidden_rdram_size, c.stream) != hipSuccess)
		return fail(RDP_MI_ERROR_DEVICE, "hidden RDRAM clear failed");
	if (hipMemsetAsync(p.d_tmem, 0, 4096, c.stream) != hipSuccess) return fail(RDP_MI_ERROR_DEVICE, "TMEM clear failed");
	if (up_log2) {
		// Sample planes zeroed, hidden planes 0x03, reference zeroed
		// (rdp_renderer.cpp:566-593).
		const size_t planes = (size_t)1 << (2 * up_log2);
		if (hipMalloc(&p.d_up_rdram, planes * rdram_size) != hipSuccess ||
		    hipMalloc(&p.d_up_hidden, planes * hidden_rdram_size) != hipSuccess ||
		    hipMalloc(&p.d_up_ref, rdram_size) != hipSuccess)
			return fail(RDP_MI_ERROR_OUT_OF_MEMORY, "upscaled domain allocation failed");
		if (hipMemsetAsync(p.d_up_rdram, 0, planes * rdram_size, c.stream) != hipSuccess ||
		    hipMemsetAsync(p.d_up_hidden, 0x03, planes * hidden_rdram_size, c.stream) != hipSuccess ||
		    hipMemsetAsync(p.d_up_ref, 0, rdram_size, c.stream) != hipSuccess)
			return fail(RDP_MI_ERROR_DEVICE, "upscaled domain clear failed");
		if (ssaa) {
			if (hipMalloc(&p.d_up_wmask, (size_t)kMaxWidth * kMaxHeight * sizeof(uint32_t)) != hipSuccess)
				return fail(RDP_MI_ERROR_OUT_OF_MEMORY, "write mask allocation failed");
			p.ssaa = true;
			p.ssaa_dither = (flags & RDP_MI_FLAG_SUPER_SAMPLED_DITHER) != 0;
			p.fe.ssaa = true;
		}
		p.up_log2 = up_log2;
		p.fe.scale_log2 = up_log2;
	}
	if (hipStreamSynchronize(c.stream) != hipSuccess) return fail(RDP_MI_ERROR_DEVICE, "device synchronisation failed");
	p.supported = true;
	// The reference's capture switch (rdp_device.cpp:60-73): failing to open
	// the file is reported and leaves capture off.
	if (const char *env = getenv("PARALLEL_RDP_DUMP_PATH")) {
		std::unique_ptr<DumpWriter> w(new DumpWriter());
		if (w->open(env, (uint32_t)rdram_size, (uint32_t)hidden_rdram_size))
			p.dump = std::move(w);
		else
			fprintf(stderr, "rdp_mi: cannot open the RDP dump %s\n", env);
	}
	*out = h;
	return 0;
}

void rdp_mi_destroy(rdp_mi_processor *h)
{
	if (!h) return;
	rdp_mi_idle(h);
	Processor &p = h->p;
	if (p.dump) p.dump->close();
	stop_ring(p);
	p.vi.release();
	(void)coherent_complete(p);
	(void)hipFree(p.d_rdram);
	(void)hipFree(p.d_hidden);
	(void)hipFree(p.d_tmem);
	if (p.up_log2) {
		(void)hipFree(p.d_up_rdram);
		(void)hipFree(p.d_up_hidden);
		(void)hipFree(p.d_up_ref);
		if (p.d_up_wmask) (void)hipFree(p.d_up_wmask);
	}
	if (p.coherent) {
		(void)hipFree(p.d_wmask);
		(void)hipHostFree(p.wb_data.ptr);
		(void)hipHostFree(p.wb_mask.ptr);
		(void)hipEventDestroy(p.wb_event);
	}
	delete h;
}

int rdp_mi_is_supported(const rdp_mi_processor *h) { return h && h->p.supported ? 1 : 0; }

void rdp_mi_enqueue_command_direct(rdp_mi_processor *h, unsigned num_words, const uint32_t *words)
{
	if (!h || !words || num_words == 0) return;
	unsigned len = Frontend::command_length(words[0]);
	if (len == 0 || len > num_words) return;
	Processor &p = h->p;
	ring_drain(p);
	std::lock_guard<std::mutex> g(p.lock);
	run_commands(p, words, len);
}

// Capture around one command (CommandProcessor::enqueue_command,
// rdp_device.cpp:874-897): the first command of a command list writes the
// RDRAM and hidden RDRAM blocks that changed since the last capture; a
// SyncFull ends the list and is recorded as SignalComplete.
static void dump_command(rdp_mi_processor *h, const uint32_t *words, unsigned len, bool before)
{
	Processor &p = h->p;
	std::lock_guard<std::mutex> g(p.dump_lock);
	if (!p.dump) return;
	if (before) {
		if (p.dump_in_list) return;
		rdp_mi_idle(h);
		if (const void *r = rdp_mi_begin_read_rdram(h)) p.dump->memory((const uint8_t *)r, p.rdram_size, false);
		if (const void *hd = rdp_mi_begin_read_hidden_rdram(h))
			p.dump->memory((const uint8_t *)hd, p.rdram_size / 2, true);
		p.dump_in_list = true;
		return;
	}
	const uint32_t op = (words[0] >> 24) & 63;
	if (op == 0x29) {
		p.dump->signal_complete();
		p.dump_in_list = false;
	} else
		p.dump->command(op, words, len);
}

void rdp_mi_enqueue_command(rdp_mi_processor *h, unsigned num_words, const uint32_t *words)
{
	if (!h || !words || num_words == 0) return;
	unsigned len = Frontend::command_length(words[0]);
	if (len == 0 || len > num_words) return;
	if (h->p.dump) dump_command(h, words, len, true);
	ring_push(h->p, words, len);
	if (h->p.dump) dump_command(h, words, len, false);
}

size_t rdp_mi_enqueue_command_list(rdp_mi_processor *h, size_t num_words, const uint32_t *words)
{
	if (!h || !words) return 0;
	if (h->p.dump) {
		// Capturing: one command at a time, each through the capture.
		size_t i = 0;
		while (i < num_words) {
			const unsigned len = Frontend::command_length(words[i]);
			if (len == 0) {
				i++;
				continue;
			}
			if (i + len > num_words) return i;
			rdp_mi_enqueue_command(h, len, words + i);
			i += len;
		}
		return i;
	}
	size_t i = 0, start = 0;
	// Hand runs of complete commands to the ring, at most kRingWords at once;
	// words with an undefined or private opcode end a run and are skipped.
	while (i < num_words) {
		unsigned len = Frontend::command_length(words[i]);
		if (len == 0 || i + len > num_words || i + len - start > kRingWords) {
			if (i > start) ring_push(h->p, words + start, i - start);
			if (len != 0 && i + len > num_words) return i;
			if (len == 0) i++;
			start = i;
			continue;
		}
		i += len;
	}
	if (i > start) ring_push(h->p, words + start, i - start);
	return i;
}

void rdp_mi_set_hold_submissions(rdp_mi_processor *h, int hold)
{
	if (!h) return;
	h->p.hold.store(hold != 0);
}

void rdp_mi_set_quirks(rdp_mi_processor *h, uint32_t quirks)
{
	if (!h) return;
	const uint32_t w[2] = { 4u << 24, quirks };
	meta_command(h->p, w, 2);
}

void rdp_mi_flush(rdp_mi_processor *h)
{
	if (!h) return;
	const uint32_t w[1] = { kMetaFlush << 24 };
	meta_command(h->p, w, 1);
}

uint64_t rdp_mi_signal_timeline(rdp_mi_processor *h)
{
	if (!h) return 0;
	Processor &p = h->p;
	const uint64_t v = ++p.timeline;
	const uint32_t w[3] = { kMetaSignal << 24, (uint32_t)v, (uint32_t)(v >> 32) };
	meta_command(p, w, 3);
	return v;
}

int rdp_mi_wait_for_timeline_timeout(rdp_mi_processor *h, uint64_t value, uint64_t timeout_us)
{
	if (!h) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "null processor");
	Processor &p = h->p;
	std::unique_lock<std::mutex> g(p.tl_lock);
	auto ready = [&] { return p.timeline_done >= value || p.timeline_error != 0; };
	if (timeout_us == UINT64_MAX)
		p.tl_cv.wait(g, ready);
	else if (!p.tl_cv.wait_for(g, std::chrono::microseconds((int64_t)std::min<uint64_t>(timeout_us, 1ull << 50)), ready))
		return set_error(RDP_MI_ERROR_TIMEOUT, "timeline value %llu not reached within %llu us",
		                 (unsigned long long)value, (unsigned long long)timeout_us);
	if (p.timeline_error)
		return set_error(p.timeline_error, "error in the processor's work: %s",
		                 p.timeline_error_msg.empty() ? "device error" : p.timeline_error_msg.c_str());
	g.unlock();
	// The caller's RDRAM gets the bytes the finished work wrote.
	if (p.coherent) return coherent_complete(p);
	return 0;
}

int rdp_mi_wait_for_timeline(rdp_mi_processor *h, uint64_t value)
{
	return rdp_mi_wait_for_timeline_timeout(h, value, UINT64_MAX);
}

uint64_t rdp_mi_completed_timeline(rdp_mi_processor *h)
{
	if (!h) return 0;
	std::lock_guard<std::mutex> g(h->p.tl_lock);
	return h->p.timeline_done;
}

int rdp_mi_idle(rdp_mi_processor *h)
{
	if (!h) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "null processor");
	return rdp_mi_wait_for_timeline(h, rdp_mi_signal_timeline(h));
}

// CommandProcessor::begin_frame_context (rdp_device.cpp:169-174): flush and
// drain the ring.
void rdp_mi_begin_frame_context(rdp_mi_processor *h)
{
	if (!h) return;
	rdp_mi_flush(h);
	ring_drain(h->p);
}

void *rdp_mi_begin_read_rdram(rdp_mi_processor *h)
{
	if (!h) return nullptr;
	Processor &p = h->p;
	// Host-coherent: the caller's memory is RDRAM (complete after a wait).
	if (p.coherent) return coherent_complete(p) ? nullptr : p.user_rdram;
	if (copy_sync(p.host_rdram(), p.d_rdram, p.rdram_size, hipMemcpyDeviceToHost)) return nullptr;
	return p.host_rdram();
}

void rdp_mi_end_write_rdram(rdp_mi_processor *h)
{
	if (!h) return;
	Processor &p = h->p;
	// Host-coherent: pages are read from the caller's memory when used.
	if (p.coherent) return;
	rdp_mi_idle(h);
	(void)copy_sync(p.d_rdram, p.host_rdram(), p.rdram_size, hipMemcpyHostToDevice);
}

void *rdp_mi_begin_read_hidden_rdram(rdp_mi_processor *h)
{
	if (!h) return nullptr;
	Processor &p = h->p;
	if (copy_sync(p.hidden_mirror.data(), p.d_hidden, p.hidden_mirror.size(), hipMemcpyDeviceToHost)) return nullptr;
	return p.hidden_mirror.data();
}

void rdp_mi_end_write_hidden_rdram(rdp_mi_processor *h)
{
	if (!h) return;
	rdp_mi_idle(h);
	Processor &p = h->p;
	(void)copy_sync(p.d_hidden, p.hidden_mirror.data(), p.hidden_mirror.size(), hipMemcpyHostToDevice);
}

size_t rdp_mi_get_rdram_size(const rdp_mi_processor *h) { return h ? h->p.rdram_size : 0; }
size_t rdp_mi_get_hidden_rdram_size(const rdp_mi_processor *h) { return h ? h->p.rdram_size / 2 : 0; }

void *rdp_mi_get_tmem(rdp_mi_processor *h)
{
	if (!h) return nullptr;
	Processor &p = h->p;
	if (copy_sync(p.tmem_mirror, p.d_tmem, 4096, hipMemcpyDeviceToHost)) return nullptr;
	return p.tmem_mirror;
}

unsigned rdp_mi_get_upscaling_factor(const rdp_mi_processor *h) { return h ? 1u << h->p.up_log2 : 0; }

static int read_upscaled(rdp_mi_processor *h, bool hidden, size_t offset, void *dst, size_t size)
{
	if (!h || (!dst && size)) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "null argument");
	Processor &p = h->p;
	if (!p.up_log2) return set_error(RDP_MI_ERROR_UNSUPPORTED, "the processor does not upscale");
	const size_t total = ((size_t)p.rdram_size >> (hidden ? 1 : 0)) << (2 * p.up_log2);
	if (offset > total || size > total - offset)
		return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "range [%zu, %zu) outside the %zu-byte planes", offset,
		                 offset + size, total);
	if (!size) return 0;
	return copy_sync(dst, (hidden ? p.d_up_hidden : p.d_up_rdram) + offset, size, hipMemcpyDeviceToHost);
}

int rdp_mi_read_upscaled_rdram(rdp_mi_processor *h, size_t offset, void *dst, size_t size)
{
	return read_upscaled(h, false, offset, dst, size);
}

int rdp_mi_read_upscaled_hidden_rdram(rdp_mi_processor *h, size_t offset, void *dst, size_t size)
{
	return read_upscaled(h, true, offset, dst, size);
}

void rdp_mi_set_vi_register(rdp_mi_processor *h, unsigned reg, uint32_t value)
{
	if (!h || reg >= RDP_MI_VI_COUNT) return;
	h->p.vi_regs[reg] = value;
	if (h->p.dump) {
		std::lock_guard<std::mutex> g(h->p.dump_lock);
		if (h->p.dump) h->p.dump->vi_register(reg, value);
	}
}

// Scanout with capture (rdp_device.cpp:1087-1092): the frame's RDRAM and
// hidden RDRAM blocks, then EndFrame.
static void dump_scanout(rdp_mi_processor *h)
{
	Processor &p = h->p;
	std::lock_guard<std::mutex> g(p.dump_lock);
	if (!p.dump) return;
	rdp_mi_idle(h);
	if (const void *r = rdp_mi_begin_read_rdram(h)) p.dump->memory((const uint8_t *)r, p.rdram_size, false);
	if (const void *hd = rdp_mi_begin_read_hidden_rdram(h)) p.dump->memory((const uint8_t *)hd, p.rdram_size / 2, true);
	p.dump->end_frame();
}

void rdp_mi_scanout_options_default(rdp_mi_scanout_options *o)
{
	if (!o) return;
	memset(o, 0, sizeof(*o));
	o->upscale_deinterlacing = 1;
	o->vi_aa = o->vi_scale = o->vi_serrate = o->vi_dither_filter = o->vi_divot_filter = o->vi_gamma_dither = 1;
}

// CommandProcessor::scanout (rdp_device.cpp:1082-1115): drain the ring, end
// the open pass and submit the processor's rendering, then queue the VI
// scanout behind it on the same stream, copying the image to `out` (host
// memory).  With `done` the copy is left in flight and `done` is recorded
// behind it; otherwise the call waits for it.
static int scanout_queue(Processor &p, const rdp_mi_scanout_options &opts, uint32_t *out, unsigned *width,
                         unsigned *height, hipEvent_t done)
{
	ring_drain(p);
	std::lock_guard<std::mutex> g(p.lock);
	if (int err = flush_now(p)) return err;
	Context &c = ctx();
	std::lock_guard<std::mutex> lg(c.lock);
	if (p.coherent) {
		// The caller may have written the frame: copy the pages the VI
		// fetches in first (resolve_coherency_external, rdp_device.cpp:1101-1106).
		if (int err = coherent_complete(p)) return err;
		uint32_t off = 0, len = 0;
		vi_memory_range(p.vi, p.vi_regs, off, len);
		std::vector<ByteRange> runs;
		page_runs({ { off, (uint64_t)off + len } }, p.rdram_size, runs);
		for (const ByteRange &r : runs)
			HIP_CHECK(hipMemcpyAsync(p.d_rdram + r.lo, p.user_rdram + r.lo, r.hi - r.lo, hipMemcpyHostToDevice, c.stream));
	}
	ViSource src;
	src.rdram_size = p.rdram_size;
	src.scale_log2 = p.up_log2;
	src.native_rdram = p.d_rdram;
	src.native_hidden = p.d_hidden;
	src.ref = p.d_up_ref;
	src.planes = p.d_up_rdram;
	src.hplanes = p.d_up_hidden;
	src.rdram = p.up_log2 ? p.d_up_rdram : p.d_rdram;
	src.hidden = p.up_log2 ? p.d_up_hidden : p.d_hidden;
	int err = vi_scanout(p.vi, p.vi_regs, src, c.d_gamma, c.stream, opts, out, width, height, done == nullptr);
	if (err) return set_error(err, "VI scanout failed");
	if (done) HIP_CHECK(hipEventRecord(done, c.stream));
	return 0;
}

int rdp_mi_scanout_sync_ex(rdp_mi_processor *h, const rdp_mi_scanout_options *opts, uint32_t *out, unsigned *width,
                           unsigned *height)
{
	if (!h || !opts || !out || !width || !height) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "null argument");
	if (h->p.dump) dump_scanout(h);
	return scanout_queue(h->p, *opts, out, width, height, nullptr);
}

int rdp_mi_scanout_sync(rdp_mi_processor *h, uint32_t *out, unsigned *width, unsigned *height)
{
	rdp_mi_scanout_options opts;
	rdp_mi_scanout_options_default(&opts);
	return rdp_mi_scanout_sync_ex(h, &opts, out, width, height);
}

int rdp_mi_scanout_buffer_create(rdp_mi_scanout_buffer **out)
{
	if (!out) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "null argument");
	*out = nullptr;
	rdp_mi_scanout_buffer *b = new (std::nothrow) rdp_mi_scanout_buffer();
	if (!b) return set_error(RDP_MI_ERROR_OUT_OF_MEMORY, "out of host memory");
	if (hipHostMalloc((void **)&b->pixels, kScanoutBufferWords * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess) {
		delete b;
		return set_error(RDP_MI_ERROR_OUT_OF_MEMORY, "scanout buffer allocation failed");
	}
	b->words = kScanoutBufferWords;
	if (hipEventCreateWithFlags(&b->fence, hipEventDisableTiming) != hipSuccess) {
		(void)hipHostFree(b->pixels);
		delete b;
		return set_error(RDP_MI_ERROR_DEVICE, "scanout fence creation failed");
	}
	*out = b;
	return 0;
}

void rdp_mi_scanout_buffer_destroy(rdp_mi_scanout_buffer *b)
{
	if (!b) return;
	if (b->in_flight) (void)hipEventSynchronize(b->fence);
	(void)hipEventDestroy(b->fence);
	(void)hipHostFree(b->pixels);
	delete b;
}

int rdp_mi_scanout_async_buffer(rdp_mi_processor *h, const rdp_mi_scanout_options *opts, rdp_mi_scanout_buffer *b)
{
	if (!h || !opts || !b) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "null argument");
	// The buffer's previous copy must land before the buffer is reused.
	if (b->in_flight) HIP_CHECK(hipEventSynchronize(b->fence));
	b->in_flight = false;
	b->width = b->height = 0;
	const size_t need = kScanoutBufferWords << (2 * h->p.up_log2);
	if (b->words < need) {
		(void)hipHostFree(b->pixels);
		b->pixels = nullptr;
		b->words = 0;
		if (hipHostMalloc((void **)&b->pixels, need * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess)
			return set_error(RDP_MI_ERROR_OUT_OF_MEMORY, "scanout buffer allocation failed");
		b->words = need;
	}
	if (h->p.dump) dump_scanout(h);
	int err = scanout_queue(h->p, *opts, b->pixels, &b->width, &b->height, b->fence);
	if (err) return err;
	b->in_flight = true;
	return 0;
}

int rdp_mi_scanout_buffer_wait(rdp_mi_scanout_buffer *b, const uint32_t **pixels, unsigned *width, unsigned *height)
{
	if (!b) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "null buffer");
	if (b->in_flight) {
		const hipError_t e = hipEventSynchronize(b->fence);
		if (e != hipSuccess) return set_error(RDP_MI_ERROR_DEVICE, "scanout fence: %s", hipGetErrorString(e));
		b->in_flight = false;
	}
	if (pixels) *pixels = b->pixels;
	if (width) *width = b->width;
	if (height) *height = b->height;
	return 0;
}

void rdp_mi_begin_vi_register_per_scanline(rdp_mi_processor *h, uint32_t flags)
{
	if (!h) return;
	std::lock_guard<std::mutex> g(h->p.lock);
	h->p.vi.begin_per_scanline(h->p.vi_regs, flags);
}

void rdp_mi_set_vi_register_for_scanline(rdp_mi_processor *h, uint32_t reg_bit, uint32_t value)
{
	if (!h) return;
	std::lock_guard<std::mutex> g(h->p.lock);
	h->p.vi.set_for_scanline(reg_bit, value);
}

void rdp_mi_latch_vi_register_for_scanline(rdp_mi_processor *h, unsigned vi_line)
{
	if (!h) return;
	std::lock_guard<std::mutex> g(h->p.lock);
	h->p.vi.latch_for_scanline(vi_line);
}

void rdp_mi_end_vi_register_per_scanline(rdp_mi_processor *h)
{
	if (!h) return;
	std::lock_guard<std::mutex> g(h->p.lock);
	h->p.vi.end_per_scanline();
}

void rdp_mi_set_validation_callback(rdp_mi_processor *h, rdp_mi_validation_fn fn, void *user)
{
	if (!h) return;
	std::lock_guard<std::mutex> g(h->p.lock);
	h->p.validation_fn = fn;
	h->p.validation_user = user;
	if (fn)
		h->p.fe.set_validation_callback([fn, user](uint32_t e, const char *m) { fn(e, m, user); });
	else
		h->p.fe.set_validation_callback(nullptr);
}

// The distinct processors of a list, their rings drained and their locks
// taken in address order (the order every multi-processor entry point uses).
struct ProcessorSet {
	std::vector<Processor *> ps;
	std::vector<std::unique_lock<std::mutex>> locks;
	int init(rdp_mi_processor *const *procs, unsigned count)
	{
		if (!procs && count) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "null processor list");
		for (unsigned i = 0; i < count; i++) {
			if (!procs[i]) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "null processor %u", i);
			if (std::find(ps.begin(), ps.end(), &procs[i]->p) == ps.end()) ps.push_back(&procs[i]->p);
		}
		for (Processor *p : ps) ring_drain(*p);
		std::vector<Processor *> order(ps);
		std::sort(order.begin(), order.end());
		for (Processor *p : order) locks.emplace_back(p->lock);
		for (Processor *p : ps) p->fe.flush();
		return 0;
	}
};

int rdp_mi_batch_flush(rdp_mi_processor *const *procs, unsigned count)
{
	ProcessorSet set;
	if (int err = set.init(procs, count)) return err;
	Context &c = ctx();
	std::lock_guard<std::mutex> g(c.lock);
	return submit_processors(set.ps.data(), (unsigned)set.ps.size(), true);
}

int rdp_mi_batch_flush_noalloc(rdp_mi_processor *const *procs, unsigned count)
{
	ProcessorSet set;
	if (int err = set.init(procs, count)) return err;
	Context &c = ctx();
	std::lock_guard<std::mutex> g(c.lock);
	uint8_t *base = g_scratch_addr[c.device];
	if (!base) return set_error(RDP_MI_ERROR_DEVICE, "the static scratch array has no device address");
	DeviceBuffer input, scr;
	input.ptr = base;
	input.size = kNoAllocInputBytes;
	input.fixed = true;
	scr.ptr = base + kNoAllocInputBytes;
	scr.size = kNoAllocScratchBytes;
	scr.fixed = true;
	PinnedBuffer staging;
	staging.ptr = g_noalloc_staging;
	staging.size = sizeof(g_noalloc_staging);
	staging.fixed = true;
	const Arenas arenas{ &input, &scr, &staging };
	int err = submit_processors(set.ps.data(), (unsigned)set.ps.size(), true, nullptr, &arenas);
	// The next call reuses the array and the host staging: finish here.
	const hipError_t e = hipStreamSynchronize(c.stream);
	if (!err && e != hipSuccess) err = set_error(RDP_MI_ERROR_DEVICE, "hipStreamSynchronize failed: %s", hipGetErrorString(e));
	return err;
}

uint64_t rdp_mi_primitive_count(const rdp_mi_processor *h) { return h ? h->p.fe.primitive_counter() : 0; }

int rdp_mi_record(rdp_mi_processor *const *procs, unsigned count, int batched, rdp_mi_recording **out)
{
	if (!out) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "null argument");
	*out = nullptr;
	ProcessorSet set;
	if (int err = set.init(procs, count)) return err;
	std::vector<Processor *> &ps = set.ps;
	for (Processor *p : ps)
		if (p->coherent)
			return set_error(RDP_MI_ERROR_UNSUPPORTED, "recordings need device-resident RDRAM (no caller RDRAM pointer)");
	Context &c = ctx();
	std::lock_guard<std::mutex> g(c.lock);
	rdp_mi_recording *h = new (std::nothrow) rdp_mi_recording();
	if (!h) return set_error(RDP_MI_ERROR_OUT_OF_MEMORY, "out of host memory");
	std::vector<std::vector<PassRef>> rounds;
	build_round_list(ps.data(), (unsigned)ps.size(), batched != 0, rounds);
	int err = prepare_rounds(c, rounds, h->r.input, h->r.scratch, h->r.launches);
	for (auto *p : ps) p->fe.completed().clear();
	if (!err && hipStreamSynchronize(c.stream) != hipSuccess) err = set_error(RDP_MI_ERROR_DEVICE, "upload failed");
	if (err) {
		delete h;
		return err;
	}
	for (auto &L : h->r.launches) {
		h->r.raster_bytes += L.raster_bytes;
		h->r.passes += L.passes;
	}
	h->r.cfg = c.cfg;
	h->r.procs = ps;
	*out = h;
	return 0;
}

static int ensure_graph(Context &c, Recording &r)
{
	if (r.exec) return 0;
	HIP_CHECK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
	int err = execute_launches(r.launches, c, r.cfg);
	hipGraph_t graph = nullptr;
	hipError_t e = hipStreamEndCapture(c.stream, &graph);
	if (err) return err;
	if (e != hipSuccess) return set_error(RDP_MI_ERROR_DEVICE, "graph capture failed: %s", hipGetErrorString(e));
	r.graph = graph;
	HIP_CHECK(hipGraphInstantiate(&r.exec, graph, nullptr, nullptr, 0));
	return 0;
}

int rdp_mi_recording_replay(rdp_mi_recording *h, unsigned iterations, int use_graph, float *iteration_ms)
{
	if (!h || iterations == 0) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "bad argument");
	Context &c = ctx();
	std::lock_guard<std::mutex> g(c.lock);
	Recording &r = h->r;
	if (use_graph) {
		int err = ensure_graph(c, r);
		if (err) return err;
	}
	std::vector<hipEvent_t> ev(iteration_ms ? iterations + 1 : 0);
	for (auto &e : ev) HIP_CHECK(hipEventCreate(&e));
	int err = 0;
	if (!ev.empty()) HIP_CHECK(hipEventRecord(ev[0], c.stream));
	for (unsigned i = 0; i < iterations && !err; i++) {
		if (use_graph) {
			hipError_t e = hipGraphLaunch(r.exec, c.stream);
			if (e != hipSuccess) err = set_error(RDP_MI_ERROR_DEVICE, "graph launch failed: %s", hipGetErrorString(e));
		} else
			err = execute_launches(r.launches, c, r.cfg);
		if (!err && !ev.empty()) HIP_CHECK(hipEventRecord(ev[i + 1], c.stream));
	}
	hipError_t se = hipStreamSynchronize(c.stream);
	if (!err && se != hipSuccess) err = set_error(RDP_MI_ERROR_DEVICE, "replay failed: %s", hipGetErrorString(se));
	if (!err && iteration_ms)
		for (unsigned i = 0; i < iterations; i++) (void)hipEventElapsedTime(&iteration_ms[i], ev[i], ev[i + 1]);
	for (auto &e : ev) (void)hipEventDestroy(e);
	return err;
}

int rdp_mi_recording_autotune(rdp_mi_recording *h, unsigned iterations, int32_t *xcd_chunk, uint32_t *min_class,
                              uint32_t *tile_size, int32_t *persistent, float *best_ms)
{
	if (!h || iterations == 0) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "bad argument");
	Context &c = ctx();
	std::lock_guard<std::mutex> g(c.lock);
	Recording &r = h->r;
	static const int32_t chunks[] = { -1, 0, 32, 64, 128 };
	std::vector<LaunchConfig> cands;
	for (bool quad : { false, true })
		for (uint32_t m = 0; m < kRasterClasses; m++)
			for (int32_t ch : chunks) cands.push_back(LaunchConfig{ ch, m, quad, false });
	// The persistent round launch takes its items in order (no XCD chunks);
	// the chunk still shapes the other launches.
	for (uint32_t m = 0; m < 2; m++) cands.push_back(LaunchConfig{ 64, m, false, true });
	std::vector<float> best(cands.size(), 1e30f);
	hipEvent_t e0, e1;
	HIP_CHECK(hipEventCreate(&e0));
	HIP_CHECK(hipEventCreate(&e1));
	int err = r.copy_state(c.stream, false);
	if (!err) err = execute_launches(r.launches, c, r.cfg);  // warm caches and code
	// Two sweeps over the candidates, best of each candidate's timed runs.
	for (int sweep = 0; sweep < 2 && !err; sweep++)
		for (size_t i = 0; i < cands.size() && !err; i++) {
			if ((err = execute_launches(r.launches, c, cands[i]))) break;
			HIP_CHECK(hipEventRecord(e0, c.stream));
			for (unsigned it = 0; it < iterations && !err; it++) err = execute_launches(r.launches, c, cands[i]);
			HIP_CHECK(hipEventRecord(e1, c.stream));
			HIP_CHECK(hipEventSynchronize(e1));
			float ms = 0;
			(void)hipEventElapsedTime(&ms, e0, e1);
			best[i] = std::min(best[i], ms / (float)iterations);
		}
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	if (!err) err = r.copy_state(c.stream, true);
	if (!err && hipStreamSynchronize(c.stream) != hipSuccess) err = set_error(RDP_MI_ERROR_DEVICE, "autotune failed");
	if (err) return err;
	size_t pick = 0;
	for (size_t i = 1; i < cands.size(); i++)
		if (best[i] < best[pick]) pick = i;
	r.cfg = cands[pick];
	r.tuned.clear();
	for (size_t i = 0; i < cands.size(); i++) r.tuned.emplace_back(cands[i], best[i]);
	// A captured graph holds the old launch shape: capture again on next use.
	if (r.exec) (void)hipGraphExecDestroy(r.exec);
	if (r.graph) (void)hipGraphDestroy(r.graph);
	r.exec = nullptr;
	r.graph = nullptr;
	if (xcd_chunk) *xcd_chunk = r.cfg.xcd_chunk;
	if (min_class) *min_class = r.cfg.min_class;
	if (tile_size) *tile_size = r.cfg.quad ? 16 : 8;
	if (persistent) *persistent = r.cfg.persistent ? 1 : 0;
	if (best_ms) *best_ms = best[pick];
	return 0;
}

unsigned rdp_mi_recording_autotune_results(const rdp_mi_recording *h, unsigned max, int32_t *xcd_chunk,
                                           uint32_t *min_class, uint32_t *tile_size, int32_t *persistent, float *ms)
{
	if (!h) return 0;
	const auto &t = h->r.tuned;
	for (size_t i = 0; i < t.size() && i < max; i++) {
		if (xcd_chunk) xcd_chunk[i] = t[i].first.xcd_chunk;
		if (min_class) min_class[i] = t[i].first.min_class;
		if (tile_size) tile_size[i] = t[i].first.quad ? 16 : 8;
		if (persistent) persistent[i] = t[i].first.persistent ? 1 : 0;
		if (ms) ms[i] = t[i].second;
	}
	return (unsigned)t.size();
}

int rdp_mi_set_launch_config(int32_t xcd_chunk, uint32_t min_class, uint32_t tile_size, int persistent)
{
	if (tile_size != 8 && tile_size != 16) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "tile size %u (8 or 16)", tile_size);
	if (persistent && tile_size != 8)
		return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "the persistent raster launch takes 8x8 tiles");
	Context &c = ctx();
	std::lock_guard<std::mutex> g(c.lock);
	c.cfg.xcd_chunk = xcd_chunk;
	c.cfg.min_class = std::min(min_class, kRasterClasses - 1);
	c.cfg.quad = tile_size == 16;
	c.cfg.persistent = persistent != 0;
	return 0;
}

int rdp_mi_recording_profile_rounds(rdp_mi_recording *h, unsigned iterations, float *round_stage_ms)
{
	if (!h || iterations == 0 || !round_stage_ms) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "bad argument");
	Context &c = ctx();
	std::lock_guard<std::mutex> g(c.lock);
	Recording &r = h->r;
	const size_t nev = r.launches.size() * 4;
	std::vector<hipEvent_t> ev(nev + 1);
	for (auto &e : ev) HIP_CHECK(hipEventCreate(&e));
	std::vector<double> acc(nev, 0.0);
	int err = r.copy_state(c.stream, false);
	for (unsigned it = 0; it < iterations && !err; it++) {
		HIP_CHECK(hipEventRecord(ev[0], c.stream));
		err = execute_launches(r.launches, c, r.cfg, ev.data() + 1);
		HIP_CHECK(hipStreamSynchronize(c.stream));
		for (size_t k = 0; k < nev && !err; k++) {
			float ms = 0;
			(void)hipEventElapsedTime(&ms, ev[k], ev[k + 1]);
			acc[k] += ms;
		}
	}
	if (!err) err = r.copy_state(c.stream, true);
	if (!err && hipStreamSynchronize(c.stream) != hipSuccess) err = set_error(RDP_MI_ERROR_DEVICE, "profile failed");
	for (auto &e : ev) (void)hipEventDestroy(e);
	for (size_t k = 0; k < nev; k++) round_stage_ms[k] = (float)(acc[k] / iterations);
	return err;
}

int rdp_mi_recording_profile_raster(rdp_mi_recording *h, unsigned iterations, unsigned max, float *ms,
                                    uint64_t *bytes, uint32_t *kind, uint32_t *tiles, unsigned *count)
{
	if (!h || iterations == 0 || !count) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "bad argument");
	Context &c = ctx();
	std::lock_guard<std::mutex> g(c.lock);
	Recording &r = h->r;
	const size_t nr = r.launches.size();
	std::vector<hipEvent_t> sev(nr * 4), rev(nr * kRasterEventsPerRound);
	for (auto &e : sev) HIP_CHECK(hipEventCreate(&e));
	for (auto &e : rev) HIP_CHECK(hipEventCreate(&e));
	std::vector<double> acc(rev.size() / 2, 0.0);
	int err = r.copy_state(c.stream, false);
	for (unsigned it = 0; it < iterations && !err; it++) {
		// Every event is recorded on each replay (launches with no work
		// still get their pair), so all pairs are valid.
		err = execute_launches(r.launches, c, r.cfg, sev.data(), rev.data());
		if (!err && hipStreamSynchronize(c.stream) != hipSuccess) err = set_error(RDP_MI_ERROR_DEVICE, "profile failed");
		for (size_t k = 0; k < acc.size() && !err; k++) {
			float t = 0;
			if (hipEventElapsedTime(&t, rev[2 * k], rev[2 * k + 1]) == hipSuccess) acc[k] += t;
		}
	}
	if (!err) err = r.copy_state(c.stream, true);
	if (!err && hipStreamSynchronize(c.stream) != hipSuccess) err = set_error(RDP_MI_ERROR_DEVICE, "profile failed");
	for (auto &e : sev) (void)hipEventDestroy(e);
	for (auto &e : rev) (void)hipEventDestroy(e);
	if (err) return err;
	// One entry per launch that ran: the fill launch, the merged light
	// class launch and each heavier class launch, in launch order.
	const uint32_t lo = std::min(r.cfg.min_class, kRasterClasses - 1);
	unsigned n = 0;
	for (size_t rd = 0; rd < nr; rd++) {
		const Launch &L = r.launches[rd];
		auto add = [&](size_t slot, uint64_t b, uint32_t k, uint32_t t) {
			if (t == 0) return;
			if (n < max) {
				if (ms) ms[n] = (float)(acc[rd * (kRasterEventsPerRound / 2) + slot] / iterations);
				if (bytes) bytes[n] = b;
				if (kind) kind[n] = k;
				if (tiles) tiles[n] = t;
			}
			n++;
		};
		add(0, L.fill_bytes, kRasterClasses, L.fill_tiles);
		uint64_t mb = 0;
		uint32_t mt = 0;
		for (uint32_t cl = 0; cl <= lo; cl++) {
			mb += L.class_bytes[cl];
			mt += L.raster_n[cl];
		}
		add(1 + lo, mb, lo, mt);
		for (uint32_t cl = lo + 1; cl < kRasterClasses; cl++) add(1 + cl, L.class_bytes[cl], cl, L.raster_n[cl]);
	}
	*count = n;
	return 0;
}

int rdp_mi_recording_profile(rdp_mi_recording *h, unsigned iterations, float stage_ms[4])
{
	if (!h || iterations == 0 || !stage_ms) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "bad argument");
	Context &c = ctx();
	std::lock_guard<std::mutex> g(c.lock);
	Recording &r = h->r;
	const size_t nev = r.launches.size() * 4;
	std::vector<hipEvent_t> ev(nev + 1);
	for (auto &e : ev) HIP_CHECK(hipEventCreate(&e));
	double acc[4] = { 0, 0, 0, 0 };
	int err = r.copy_state(c.stream, false);
	for (unsigned it = 0; it < iterations && !err; it++) {
		HIP_CHECK(hipEventRecord(ev[0], c.stream));
		err = execute_launches(r.launches, c, r.cfg, ev.data() + 1);
		HIP_CHECK(hipStreamSynchronize(c.stream));
		for (size_t k = 0; k < nev && !err; k++) {
			float ms = 0;
			(void)hipEventElapsedTime(&ms, ev[k], ev[k + 1]);
			acc[k % 4] += ms;
		}
	}
	if (!err) err = r.copy_state(c.stream, true);
	if (!err && hipStreamSynchronize(c.stream) != hipSuccess) err = set_error(RDP_MI_ERROR_DEVICE, "profile failed");
	for (auto &e : ev) (void)hipEventDestroy(e);
	for (int k = 0; k < 4; k++) stage_ms[k] = (float)(acc[k] / iterations);
	return err;
}

uint64_t rdp_mi_recording_raster_bytes(const rdp_mi_recording *h) { return h ? h->r.raster_bytes : 0; }

unsigned rdp_mi_recording_rounds(const rdp_mi_recording *h) { return h ? (unsigned)h->r.launches.size() : 0; }

unsigned rdp_mi_recording_passes(const rdp_mi_recording *h) { return h ? h->r.passes : 0; }

unsigned rdp_mi_recording_raster_launches(const rdp_mi_recording *h)
{
	if (!h) return 0;
	const LaunchConfig &cfg = h->r.cfg;
	const uint32_t lo = std::min(cfg.min_class, kRasterClasses - 1);
	unsigned n = 0;
	for (const Launch &L : h->r.launches) {
		uint32_t merged = 0;
		for (uint32_t c = 0; c <= lo; c++) merged += L.raster_n[c];
		n += merged != 0;
		n += L.fill_n != 0;
		for (uint32_t c = lo + 1; c < kRasterClasses; c++) n += L.raster_n[c] != 0;
	}
	return n;
}

int rdp_mi_server_create(rdp_mi_processor *const *procs, unsigned count, size_t hbm_budget_bytes, unsigned max_batch,
                         unsigned decode_threads, unsigned streams, rdp_mi_server **out)
{
	if (!out || !procs || !count) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "bad argument");
	*out = nullptr;
	rdp_mi_server *h = new (std::nothrow) rdp_mi_server();
	if (!h) return set_error(RDP_MI_ERROR_OUT_OF_MEMORY, "out of host memory");
	Server &s = h->s;
	for (unsigned i = 0; i < count; i++) {
		if (!procs[i]) {
			delete h;
			return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "null processor %u", i);
		}
		if (procs[i]->p.coherent) {
			delete h;
			return set_error(RDP_MI_ERROR_UNSUPPORTED, "serving needs device-resident RDRAM (processor %u has a caller RDRAM pointer)", i);
		}
		s.procs.push_back(&procs[i]->p);
	}
	s.budget = hbm_budget_bytes ? hbm_budget_bytes : ((size_t)1 << 32);
	s.max_batch = max_batch ? max_batch : count;
	// Defaults: a quarter of the host threads (at most 8) decode, leaving the
	// rest to the scheduler, submitter and completion threads and the
	// caller (more decode threads lengthened the latency tail on a 16-thread
	// host); three streams.
	s.decode_threads = decode_threads ? decode_threads
	                                  : std::max(1u, std::min(8u, std::thread::hardware_concurrency() / 4));
	for (unsigned i = 0; i < (streams ? std::min(streams, 8u) : 3u); i++)
		s.lanes.emplace_back(new Server::LaneState());
	s.queues.resize(count);
	s.busy.assign(count, false);
	*out = h;
	return 0;
}

int rdp_mi_server_submit(rdp_mi_server *h, unsigned instance, const uint32_t *words, size_t num_words,
                         double arrival_us, uint64_t *ticket)
{
	if (!h || !words || instance >= h->s.procs.size()) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "bad argument");
	Server &s = h->s;
	std::lock_guard<std::mutex> g(s.m);
	if (s.stopping) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "server is draining");
	ServeRequest r;
	r.ticket = s.next_ticket++;
	r.instance = instance;
	r.arrival_us = s.started ? std::max(arrival_us, s.now_us()) : arrival_us;
	r.words.assign(words, words + num_words);
	if (ticket) *ticket = r.ticket;
	// Keep each processor's queue in arrival order.
	auto &q = s.queues[instance];
	auto it = q.end();
	while (it != q.begin() && (it - 1)->arrival_us > r.arrival_us) --it;
	q.insert(it, std::move(r));
	s.outstanding++;
	s.cv.notify_all();
	return 0;
}

int rdp_mi_server_start(rdp_mi_server *h)
{
	if (!h) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "null server");
	Server &s = h->s;
	std::lock_guard<std::mutex> g(s.m);
	if (s.started) return 0;
	Context &c = ctx();
	{
		std::lock_guard<std::mutex> lk(c.lock);
		int err = c.init();
		if (err) return err;
		// Work queued on the context stream for these processors finishes
		// before the server's lanes touch them.
		if (hipStreamSynchronize(c.stream) != hipSuccess) return set_error(RDP_MI_ERROR_DEVICE, "synchronisation failed");
	}
	// Size each lane's arenas up front (growing one later waits for the
	// device): packed inputs run at about a quarter of a batch's estimate.
	const size_t scratch_bytes = std::min<size_t>(s.budget, (size_t)256 << 20);
	const size_t input_bytes = scratch_bytes / 4;
	for (auto &l : s.lanes) {
		Lane &ln = l->lane;
		if (int err = ln.create()) return err;
		if (ln.staging.reserve(input_bytes) || ln.input.reserve(input_bytes, ln.stream) ||
		    ln.scratch.reserve(scratch_bytes, ln.stream))
			return set_error(RDP_MI_ERROR_OUT_OF_MEMORY, "lane arenas of %zu bytes failed", input_bytes * 2 + scratch_bytes);
		// First use of a stream or event costs far more than later ones:
		// create the per-round events and touch every stream now.
		if (int err = ln.ensure_round_events(4)) return err;
		HIP_CHECK(hipEventCreateWithFlags(&l->done_ev, hipEventDisableTiming));
		if (s.digests) {
			const size_t n = s.max_batch;
			HIP_CHECK(hipHostMalloc((void **)&l->h_targets, n * sizeof(DigestTarget), hipHostMallocDefault));
			HIP_CHECK(hipMalloc((void **)&l->d_targets, n * sizeof(DigestTarget)));
			HIP_CHECK(hipHostMalloc((void **)&l->h_digests, n * sizeof(uint64_t), hipHostMallocDefault));
			HIP_CHECK(hipMalloc((void **)&l->d_digests, n * sizeof(uint64_t)));
		}
		for (hipStream_t st : { ln.stream, ln.side, ln.prep }) {
			HIP_CHECK(hipEventRecord(l->done_ev, st));
			HIP_CHECK(hipEventSynchronize(l->done_ev));
		}
	}
	s.pool.reset(new ThreadPool(s.decode_threads));
	s.t0 = std::chrono::steady_clock::now();
	s.started = true;
	s.scheduler = std::thread([&s] { s.scheduler_loop(); });
	for (auto &l : s.lanes) {
		Server::LaneState *ls = l.get();
		ls->submitter = std::thread([&s, ls] { s.submitter_loop(ls); });
		ls->completer = std::thread([&s, ls] { s.completion_loop(*ls); });
	}
	return 0;
}

int rdp_mi_server_drain(rdp_mi_server *h)
{
	if (!h) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "null server");
	Server &s = h->s;
	{
		std::unique_lock<std::mutex> g(s.m);
		if (!s.started) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "server not started");
		s.cv.wait(g, [&s] { return s.outstanding == 0; });
		if (s.error) return set_error(s.error, "%s", s.error_msg.c_str());
	}
	return 0;
}

int rdp_mi_server_reset(rdp_mi_server *h, double start_in_us)
{
	if (!h) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "null server");
	Server &s = h->s;
	std::lock_guard<std::mutex> g(s.m);
	if (!s.started) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "server not started");
	if (s.outstanding != 0) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "server has requests in progress");
	s.results.clear();
	s.n_batches = s.n_frames = s.max_frames = s.max_est_bytes = s.max_arena_bytes = 0;
	s.submit_time_us = s.max_submit_us = 0;
	s.t0 = std::chrono::steady_clock::now() + std::chrono::microseconds((int64_t)std::max(0.0, start_in_us));
	s.cv.notify_all();
	return 0;
}

size_t rdp_mi_server_results(rdp_mi_server *h, uint64_t *tickets, uint32_t *instances, double *arrival_us,
                             double *admit_us, double *submit_us, double *complete_us, size_t max)
{
	if (!h) return 0;
	Server &s = h->s;
	std::lock_guard<std::mutex> g(s.m);
	const size_t n = s.results.size();
	for (size_t i = 0; i < n && i < max; i++) {
		const ServeResult &r = s.results[i];
		if (tickets) tickets[i] = r.ticket;
		if (instances) instances[i] = r.instance;
		if (arrival_us) arrival_us[i] = r.arrival_us;
		if (admit_us) admit_us[i] = r.admit_us;
		if (submit_us) submit_us[i] = r.submit_us;
		if (complete_us) complete_us[i] = r.complete_us;
	}
	return n;
}

int rdp_mi_server_set_batch_window(rdp_mi_server *h, double window_us)
{
	if (!h || !(window_us >= 0)) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "bad argument");
	std::lock_guard<std::mutex> g(h->s.m);
	h->s.batch_window_us = window_us;
	h->s.submit_cv.notify_all();
	return 0;
}

int rdp_mi_server_set_digests(rdp_mi_server *h, int enable)
{
	if (!h) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "null server");
	Server &s = h->s;
	std::lock_guard<std::mutex> g(s.m);
	if (s.started) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "server already started");
	s.digests = enable != 0;
	return 0;
}

size_t rdp_mi_server_digests(rdp_mi_server *h, uint64_t *tickets, uint64_t *digests, int32_t *status, size_t max)
{
	if (!h) return 0;
	Server &s = h->s;
	std::lock_guard<std::mutex> g(s.m);
	const size_t n = s.results.size();
	for (size_t i = 0; i < n && i < max; i++) {
		const ServeResult &r = s.results[i];
		if (tickets) tickets[i] = r.ticket;
		if (digests) digests[i] = r.digest;
		if (status) status[i] = r.status;
	}
	return n;
}

int rdp_mi_state_digest(rdp_mi_processor *h, uint64_t *digest)
{
	if (!h || !digest) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "null argument");
	Processor &p = h->p;
	Context &c = ctx();
	std::lock_guard<std::mutex> g(c.lock);
	if (!c.d_digest) {
		HIP_CHECK(hipMalloc((void **)&c.d_digest, sizeof(DigestTarget) + 256));
	}
	const DigestTarget t = digest_target(p);
	DigestTarget *d_t = reinterpret_cast<DigestTarget *>(c.d_digest + 256);
	uint64_t *d_out = reinterpret_cast<uint64_t *>(c.d_digest);
	HIP_CHECK(hipMemcpyAsync(d_t, &t, sizeof(t), hipMemcpyHostToDevice, c.stream));
	HIP_CHECK(launch_digest(d_t, 1, digest_words(p), d_out, c.stream));
	HIP_CHECK(hipMemcpyAsync(digest, d_out, sizeof(uint64_t), hipMemcpyDeviceToHost, c.stream));
	HIP_CHECK(hipStreamSynchronize(c.stream));
	return 0;
}

void rdp_mi_server_stats(rdp_mi_server *h, uint64_t out[7])
{
	if (!h || !out) return;
	Server &s = h->s;
	std::lock_guard<std::mutex> g(s.m);
	out[0] = s.n_batches;
	out[1] = s.n_frames;
	out[2] = s.max_frames;
	out[3] = s.max_est_bytes;
	out[4] = s.max_arena_bytes;
	out[5] = (uint64_t)s.submit_time_us;
	out[6] = (uint64_t)s.max_submit_us;
}

void rdp_mi_server_destroy(rdp_mi_server *h)
{
	if (!h) return;
	Server &s = h->s;
	{
		// Requests still queued are rendered before the threads end.
		std::lock_guard<std::mutex> g(s.m);
		s.stopping = true;
		s.cv.notify_all();
		s.submit_cv.notify_all();
		for (auto &l : s.lanes) l->cv.notify_all();
	}
	if (s.scheduler.joinable()) s.scheduler.join();
	for (auto &l : s.lanes) {
		if (l->submitter.joinable()) l->submitter.join();
		if (l->completer.joinable()) l->completer.join();
	}
	s.pool.reset();
	for (auto &l : s.lanes) {
		if (l->done_ev) (void)hipEventDestroy(l->done_ev);
		l->lane.release();
		if (l->h_targets) (void)hipHostFree(l->h_targets);
		if (l->h_digests) (void)hipHostFree(l->h_digests);
		if (l->d_targets) (void)hipFree(l->d_targets);
		if (l->d_digests) (void)hipFree(l->d_digests);
	}
	delete h;
}

void rdp_mi_recording_destroy(rdp_mi_recording *h)
{
	if (!h) return;
	Context &c = ctx();
	std::lock_guard<std::mutex> g(c.lock);
	(void)hipStreamSynchronize(c.stream);
	if (h->r.exec) (void)hipGraphExecDestroy(h->r.exec);
	if (h->r.graph) (void)hipGraphDestroy(h->r.graph);
	if (h->r.input.ptr) (void)hipFree(h->r.input.ptr);
	if (h->r.scratch.ptr) (void)hipFree(h->r.scratch.ptr);
	if (h->r.saved.ptr) (void)hipFree(h->r.saved.ptr);
	delete h;
}

int rdp_mi_framebuffer_view(rdp_mi_processor *h, uint32_t mode, uint32_t fb_addr, uint32_t fb_size,
                            uint32_t depth_addr, uint32_t width, uint32_t height, uint32_t *out)
{
	if (!h || !out || mode > RDP_MI_VIEW_COVERAGE || width == 0 || width > 4096 || height > 4096)
		return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "bad argument");
	if (height == 0) height = std::max(1u, width * 3 / 4);
	Processor &p = h->p;
	ring_drain(p);
	std::lock_guard<std::mutex> g(p.lock);
	if (int err = flush_now(p)) return err;
	Context &c = ctx();
	std::lock_guard<std::mutex> lg(c.lock);
	const size_t bytes = (size_t)width * height * 4;
	DeviceBuffer &buf = c.view;
	if (buf.reserve(bytes, c.stream)) return set_error(RDP_MI_ERROR_OUT_OF_MEMORY, "view buffer");
	if (p.coherent) {
		// The views read the device copy: bring in the caller's pages.
		if (int err = coherent_complete(p)) return err;
		HIP_CHECK(hipMemcpyAsync(p.d_rdram, p.user_rdram, p.rdram_size, hipMemcpyHostToDevice, c.stream));
	}
	ViewArgs a{ p.d_rdram, p.d_hidden, p.rdram_size, mode, fb_addr, fb_size, depth_addr, width, height };
	HIP_CHECK(launch_view(a, reinterpret_cast<uint32_t *>(buf.ptr), c.stream));
	HIP_CHECK(hipMemcpyAsync(out, buf.ptr, bytes, hipMemcpyDeviceToHost, c.stream));
	HIP_CHECK(hipStreamSynchronize(c.stream));
	return 0;
}

// ---- RDPDUMP2 capture and replay ----

struct rdp_mi_dump_writer {
	DumpWriter w;
};

int rdp_mi_dump_writer_open(const char *path, uint32_t rdram_size, uint32_t hidden_rdram_size, rdp_mi_dump_writer **out)
{
	if (!out || !path) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "null argument");
	*out = nullptr;
	rdp_mi_dump_writer *w = new (std::nothrow) rdp_mi_dump_writer();
	if (!w) return set_error(RDP_MI_ERROR_OUT_OF_MEMORY, "out of host memory");
	if (!w->w.open(path, rdram_size, hidden_rdram_size)) {
		delete w;
		return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "cannot create %s", path);
	}
	*out = w;
	return 0;
}

void rdp_mi_dump_writer_flush_rdram(rdp_mi_dump_writer *w, const void *rdram, uint32_t size)
{
	if (w && rdram) w->w.memory((const uint8_t *)rdram, size, false);
}

void rdp_mi_dump_writer_flush_hidden_rdram(rdp_mi_dump_writer *w, const void *hidden, uint32_t size)
{
	if (w && hidden) w->w.memory((const uint8_t *)hidden, size, true);
}

void rdp_mi_dump_writer_command(rdp_mi_dump_writer *w, uint32_t op, const uint32_t *words, uint32_t num_words)
{
	if (w && (words || !num_words)) w->w.command(op, words, num_words);
}

void rdp_mi_dump_writer_set_vi_register(rdp_mi_dump_writer *w, uint32_t reg, uint32_t value)
{
	if (w) w->w.vi_register(reg, value);
}

void rdp_mi_dump_writer_end_frame(rdp_mi_dump_writer *w)
{
	if (w) w->w.end_frame();
}

void rdp_mi_dump_writer_signal_complete(rdp_mi_dump_writer *w)
{
	if (w) w->w.signal_complete();
}

void rdp_mi_dump_writer_close(rdp_mi_dump_writer *w)
{
	if (!w) return;
	w->w.close();
	delete w;
}

int rdp_mi_begin_dump(rdp_mi_processor *h, const char *path)
{
	if (!h || !path) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "null argument");
	Processor &p = h->p;
	ring_drain(p);
	std::lock_guard<std::mutex> g(p.dump_lock);
	if (p.dump) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "the processor is already capturing");
	std::unique_ptr<DumpWriter> w(new (std::nothrow) DumpWriter());
	if (!w) return set_error(RDP_MI_ERROR_OUT_OF_MEMORY, "out of host memory");
	if (!w->open(path, p.rdram_size, p.rdram_size / 2)) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "cannot create %s", path);
	p.dump = std::move(w);
	p.dump_in_list = false;
	return 0;
}

void rdp_mi_end_dump(rdp_mi_processor *h)
{
	if (!h) return;
	Processor &p = h->p;
	ring_drain(p);
	std::lock_guard<std::mutex> g(p.dump_lock);
	if (p.dump) p.dump->close();
	p.dump.reset();
	p.dump_in_list = false;
}

struct rdp_mi_dump_player {
	DumpReader r;
};

int rdp_mi_dump_player_open(const char *path, rdp_mi_dump_player **out)
{
	if (!out || !path) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "null argument");
	*out = nullptr;
	rdp_mi_dump_player *pl = new (std::nothrow) rdp_mi_dump_player();
	if (!pl) return set_error(RDP_MI_ERROR_OUT_OF_MEMORY, "out of host memory");
	if (!pl->r.open(path)) {
		delete pl;
		return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "%s is not an RDPDUMP2 capture", path);
	}
	*out = pl;
	return 0;
}

uint32_t rdp_mi_dump_player_rdram_size(const rdp_mi_dump_player *pl) { return pl ? pl->r.rdram_size() : 0; }
uint32_t rdp_mi_dump_player_hidden_rdram_size(const rdp_mi_dump_player *pl) { return pl ? pl->r.hidden_size() : 0; }

int rdp_mi_dump_player_rewind(rdp_mi_dump_player *pl)
{
	if (!pl) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "null player");
	return pl->r.rewind() ? 0 : set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "rewind failed");
}

int rdp_mi_dump_player_next(rdp_mi_dump_player *pl, rdp_mi_dump_event *ev)
{
	if (!pl || !ev) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "null argument");
	DumpEvent e;
	const int r = pl->r.next(e);
	memset(ev, 0, sizeof(*ev));
	if (r < 0) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "malformed capture record");
	if (r == 0) {
		ev->kind = RDP_MI_DUMP_EOF;
		return 0;
	}
	ev->kind = e.kind;
	ev->op = e.op;
	ev->num_words = e.num_words;
	ev->words = e.words;
	ev->vi_register = e.vi_register;
	ev->vi_value = e.vi_value;
	ev->offset = e.offset;
	ev->size = e.size;
	return 1;
}

const uint8_t *rdp_mi_dump_player_rdram(const rdp_mi_dump_player *pl) { return pl ? pl->r.rdram() : nullptr; }
const uint8_t *rdp_mi_dump_player_hidden_rdram(const rdp_mi_dump_player *pl) { return pl ? pl->r.hidden() : nullptr; }

// The reference's replay driver (replayer_driver_parallel.cpp:124-205).
int rdp_mi_dump_player_apply(rdp_mi_dump_player *pl, const rdp_mi_dump_event *ev, rdp_mi_processor *h)
{
	if (!pl || !ev || !h) return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "null argument");
	switch (ev->kind) {
	case RDP_MI_DUMP_RDP_COMMAND:
		if (ev->num_words) rdp_mi_enqueue_command(h, ev->num_words, ev->words);
		return 0;
	case RDP_MI_DUMP_SET_VI_REGISTER:
		if (ev->vi_register >= RDP_MI_VI_COUNT)
			return set_error(RDP_MI_ERROR_INVALID_ARGUMENT, "VI register %u", ev->vi_register);
		rdp_mi_set_vi_register(h, ev->vi_register, ev->vi_value);
		return 0;
	case RDP_MI_DUMP_SIGNAL_COMPLETE:
		rdp_mi_flush(h);
		return 0;
	case RDP_MI_DUMP_UPDATE_DRAM_FLUSH:
	case RDP_MI_DUMP_UPDATE_HIDDEN_DRAM_FLUSH: {
		const bool hidden = ev->kind == RDP_MI_DUMP_UPDATE_HIDDEN_DRAM_FLUSH;
		if (int err = rdp_mi_idle(h)) return err;
		void *dst = hidden ? rdp_mi_begin_read_hidden_rdram(h) : rdp_mi_begin_read_rdram(h);
		if (!dst) return set_error(RDP_MI_ERROR_DEVICE, "RDRAM read-back failed");
		const size_t n = hidden ? std::min<size_t>(pl->r.hidden_size(), rdp_mi_get_hidden_rdram_size(h))
		                        : std::min<size_t>(pl->r.rdram_size(), rdp_mi_get_rdram_size(h));
		memcpy(dst, hidden ? pl->r.hidden() : pl->r.rdram(), n);
		if (hidden)
			rdp_mi_end_write_hidden_rdram(h);
		else
			rdp_mi_end_write_rdram(h);
		return 0;
	}
	default: return 0;  // block updates land in the shadow; EndFrame and EOF are the caller's
	}
}

void rdp_mi_dump_player_destroy(rdp_mi_dump_player *pl) { delete pl; }

}  // extern "C"


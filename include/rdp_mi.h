/*
 * rdp_mi.h - C ABI of the MI355X RDP rendering library (librdp_mi.so).
 *
 * Drop-in boundary for the reference's RDP::CommandProcessor
 * (parallel-rdp/rdp_device.hpp:111-198).  One opaque processor per emulated
 * RDP; plain pointers and sizes only.  Each entry point cites the reference
 * interface it replaces.
 *
 * Memory model: RDRAM, hidden RDRAM and TMEM stay resident in GPU memory.
 * rdp_mi_begin_read_* copies the device state into a host mirror (call it
 * after rdp_mi_idle or rdp_mi_wait_for_timeline); rdp_mi_end_write_* uploads
 * the host mirror back.  When the caller passes its own RDRAM pointer at
 * creation, that buffer is the host mirror.
 *
 * Errors: entry points returning int return 0 on success and a negative
 * RDP_MI_ERROR_* code otherwise; rdp_mi_last_error() gives a message.
 * Nothing throws across the ABI.
 */
#ifndef RDP_MI_H
#define RDP_MI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rdp_mi_processor rdp_mi_processor;

enum {
	RDP_MI_OK = 0,
	RDP_MI_ERROR_INVALID_ARGUMENT = -1,
	RDP_MI_ERROR_OUT_OF_MEMORY = -2,
	RDP_MI_ERROR_DEVICE = -3,
	RDP_MI_ERROR_UNSUPPORTED = -4,
	RDP_MI_ERROR_TIMEOUT = -5,
	RDP_MI_ERROR_LIMIT = -6,
};

/* Flags accepted by rdp_mi_create (rdp_device.hpp:43-51).  One upscaling
 * flag renders every render pass whose colour image is RGBA5551 or RGBA8888
 * and at least 256 pixels wide a second time at 2, 4 or 8 times the
 * resolution into factor^2 sample planes, and the VI scans out from those
 * planes at factor times the size (rdp_device.cpp:188-203, rdp_renderer.cpp:
 * 545-601, 2223-2342, video_interface.cpp:1215-1318).  Native RDRAM is
 * rendered exactly as without the flag.  Not with host-coherent RDRAM
 * (RDP_MI_ERROR_UNSUPPORTED).  With RDP_MI_FLAG_SUPER_SAMPLED_READ_BACK as
 * well, every pass renders at scale only and is resolved into native RDRAM
 * (the mean of each written pixel's samples, depth from its first sample,
 * hidden RDRAM untouched; RDP_MI_FLAG_SUPER_SAMPLED_DITHER adds 4x4 ordered
 * dither to 16-bit colour; update_upscaled_domain_resolve.comp); without an
 * upscaling flag it is rejected with RDP_MI_ERROR_UNSUPPORTED.  The VI's
 * downscale / upscale-deinterlace blits (float presentation) are not
 * offered. */
enum {
	RDP_MI_FLAG_HOST_VISIBLE_HIDDEN_RDRAM = 1 << 0,
	RDP_MI_FLAG_HOST_VISIBLE_TMEM = 1 << 1,
	RDP_MI_FLAG_UPSCALING_2X = 1 << 2,
	RDP_MI_FLAG_UPSCALING_4X = 1 << 3,
	RDP_MI_FLAG_UPSCALING_8X = 1 << 4,
	RDP_MI_FLAG_SUPER_SAMPLED_READ_BACK = 1 << 5,
	RDP_MI_FLAG_SUPER_SAMPLED_DITHER = 1 << 6,
};

/* VI registers (rdp_common.hpp:324-341). */
enum {
	RDP_MI_VI_CONTROL = 0, RDP_MI_VI_ORIGIN, RDP_MI_VI_WIDTH, RDP_MI_VI_INTR, RDP_MI_VI_V_CURRENT_LINE,
	RDP_MI_VI_TIMING, RDP_MI_VI_V_SYNC, RDP_MI_VI_H_SYNC, RDP_MI_VI_LEAP, RDP_MI_VI_H_START, RDP_MI_VI_V_START,
	RDP_MI_VI_V_BURST, RDP_MI_VI_X_SCALE, RDP_MI_VI_Y_SCALE, RDP_MI_VI_COUNT
};

/* CommandProcessor::CommandProcessor(device, rdram_ptr, rdram_offset,
 * rdram_size, hidden_rdram_size, flags), rdp_device.cpp:44-162.
 * rdram_size must be a power of two; hidden_rdram_size must be
 * rdram_size / 2.  Hidden RDRAM starts at 0x03 per byte, TMEM at zero.
 *
 * Without rdram_ptr, RDRAM lives on the device (zeroed) and the host sees it
 * through rdp_mi_begin_read_rdram (a copy, after idle or a wait) and
 * rdp_mi_end_write_rdram (uploads the host copy).
 *
 * With rdram_ptr, the caller's memory at rdram_ptr + rdram_offset is RDRAM,
 * as with the reference's host-memory RDRAM when the device cannot import it
 * (rdp_device.cpp:75-112, rdp_renderer.cpp:2523-2907): the device keeps a
 * copy of it; each submission first copies in the 4 KiB pages its render
 * passes read (colour and depth image rows, TMEM upload sources), and a
 * scanout the pages the VI fetches (video_interface.cpp:369-396); the raster
 * marks every byte it writes, and once the work is complete (a wait,
 * idle) only those bytes are written into the caller's memory
 * (masked_rdram_resolve.comp:47-70), so caller writes elsewhere in the same
 * pages are kept.  The caller may write RDRAM at any time between
 * submissions without rdp_mi_end_write_rdram (a no-op in this mode);
 * rdp_mi_begin_read_rdram returns rdram_ptr + rdram_offset.  Recordings and
 * servers need device-resident RDRAM and reject such processors. */
int rdp_mi_create(rdp_mi_processor **out, void *rdram_ptr, size_t rdram_offset, size_t rdram_size,
                  size_t hidden_rdram_size, uint32_t flags);

/* CommandProcessor::~CommandProcessor, rdp_device.cpp:164-167 (implies idle). */
void rdp_mi_destroy(rdp_mi_processor *p);

/* CommandProcessor::device_is_supported, rdp_device.cpp:225-228. */
int rdp_mi_is_supported(const rdp_mi_processor *p);

/* CommandProcessor::enqueue_command, rdp_device.cpp:874-897, through the
 * CommandRing (command_ring.cpp:71-134): the words are copied into the
 * processor's command ring and decoded on its own thread, so the call returns
 * without decoding (it blocks only while the ring is full).  words holds one
 * complete command.  Opcodes 1-3 are private meta commands and are dropped. */
void rdp_mi_enqueue_command(rdp_mi_processor *p, unsigned num_words, const uint32_t *words);

/* CommandProcessor::enqueue_command_direct, rdp_device.cpp:899-959: decodes
 * on the calling thread (after the ring has drained, so the order of
 * commands is kept). */
void rdp_mi_enqueue_command_direct(rdp_mi_processor *p, unsigned num_words, const uint32_t *words);

/* Enqueue a packed run of complete commands through the command ring (no
 * reference counterpart: the reference's callers loop over enqueue_command,
 * which this performs with one ring hand-off per 64 Ki words).  Returns the
 * number of words consumed; a truncated trailing command is left unconsumed,
 * a word with an undefined or private opcode is skipped. */
size_t rdp_mi_enqueue_command_list(rdp_mi_processor *p, size_t num_words, const uint32_t *words);

/* No reference counterpart (a switch for the batched and recorded
 * submissions below).  With hold != 0 the processor's completed render
 * passes stay on the host: the submissions the reference makes on its own
 * (after 500 us without commands, command_ring.cpp:103-121, and when many
 * passes queue up) are suspended, so rdp_mi_record and rdp_mi_batch_flush
 * take every pass enqueued since the last submission.  Explicit flush,
 * signal_timeline, idle and scanout still submit.  Host memory grows with
 * the held passes.  Default: 0 (the reference's behaviour). */
void rdp_mi_set_hold_submissions(rdp_mi_processor *p, int hold);

/* CommandProcessor::set_quirks, rdp_device.cpp:961-968 (in-band). */
void rdp_mi_set_quirks(rdp_mi_processor *p, uint32_t quirks);

/* CommandProcessor::flush / signal_timeline / wait_for_timeline / idle /
 * begin_frame_context, rdp_device.cpp:169-174, 1039-1080.
 * flush and signal_timeline are in-band: they travel through the command
 * ring behind the commands enqueued before them.  signal_timeline returns the
 * next value of the processor's monotonic timeline; a completion thread
 * (worker_thread.hpp:39-126) publishes it once the device has finished every
 * command enqueued before the signal.  wait_for_timeline blocks the calling
 * thread until the value is published, also for a value not yet signalled
 * (worker_thread.hpp:69-73): it returns once another thread's signal_timeline
 * reaches it.  It returns RDP_MI_ERROR_DEVICE if the device reported an error
 * for the processor's work. */
void rdp_mi_flush(rdp_mi_processor *p);
uint64_t rdp_mi_signal_timeline(rdp_mi_processor *p);
int rdp_mi_wait_for_timeline(rdp_mi_processor *p, uint64_t value);
/* wait_for_timeline with a time limit (no reference counterpart): returns
 * RDP_MI_ERROR_TIMEOUT when the value is not published within timeout_us. */
int rdp_mi_wait_for_timeline_timeout(rdp_mi_processor *p, uint64_t value, uint64_t timeout_us);
/* The last timeline value whose work is complete (no reference counterpart;
 * the reference keeps it in thread_timeline_value, rdp_device.cpp:1184-1188). */
uint64_t rdp_mi_completed_timeline(rdp_mi_processor *p);
int rdp_mi_idle(rdp_mi_processor *p);
void rdp_mi_begin_frame_context(rdp_mi_processor *p);

/* Memory interaction, rdp_device.cpp:997-1037. */
void *rdp_mi_begin_read_rdram(rdp_mi_processor *p);
void rdp_mi_end_write_rdram(rdp_mi_processor *p);
void *rdp_mi_begin_read_hidden_rdram(rdp_mi_processor *p);
void rdp_mi_end_write_hidden_rdram(rdp_mi_processor *p);
size_t rdp_mi_get_rdram_size(const rdp_mi_processor *p);
size_t rdp_mi_get_hidden_rdram_size(const rdp_mi_processor *p);
void *rdp_mi_get_tmem(rdp_mi_processor *p);

/* Renderer::get_scaling_factor (rdp_renderer.hpp): 1, or the factor of the
 * processor's upscaling flag. */
unsigned rdp_mi_get_upscaling_factor(const rdp_mi_processor *p);

/* Copies of the upscaled domain (after idle or a wait, as with
 * rdp_mi_begin_read_rdram).  The reference keeps these buffers internal to
 * its renderer (upscaling_multisampled_rdram / _hidden_rdram,
 * rdp_renderer.cpp:566-593); they are exported here for inspection and
 * tests.  The RDRAM planes are factor^2 arrays of rdram_size bytes laid out
 * like RDRAM, plane sy * factor + sx holding sample (sx, sy) of each pixel;
 * the hidden planes are factor^2 arrays of rdram_size / 2 bytes.  size bytes
 * from byte offset into the concatenated planes.  Returns 0,
 * RDP_MI_ERROR_UNSUPPORTED for a processor without upscaling or
 * RDP_MI_ERROR_INVALID_ARGUMENT for a range outside the planes. */
int rdp_mi_read_upscaled_rdram(rdp_mi_processor *p, size_t offset, void *dst, size_t size);
int rdp_mi_read_upscaled_hidden_rdram(rdp_mi_processor *p, size_t offset, void *dst, size_t size);

/* CommandProcessor::set_vi_register, rdp_device.cpp:970-974. */
void rdp_mi_set_vi_register(rdp_mi_processor *p, unsigned reg, uint32_t value);

/* CommandProcessor::scanout_sync, rdp_device.cpp:1159-1182: drains work,
 * scans out through the VI and writes width*height RGBA8 words (alpha 255)
 * into out, which must hold 640 * 576 words (times factor^2 for an
 * upscaling processor, whose image is factor times larger in each
 * direction).  A blank or invalid frame gives 0x0. */
int rdp_mi_scanout_sync(rdp_mi_processor *p, uint32_t *out, unsigned *width, unsigned *height);

/* ScanoutOptions (video_interface.hpp:31-80) without upscaling, downscaling
 * or external memory.  rdp_mi_scanout_options_default gives the reference's
 * defaults: no crop, no persistence, no blending with the previous frame,
 * upscale deinterlacing (each field scanned out progressively, 240 or 288
 * lines) and every VI filter enabled.  The reference's conformance driver
 * uses blend_previous_frame = 1, upscale_deinterlacing = 0
 * (replayer_driver_parallel.cpp:194-206). */
typedef struct rdp_mi_scanout_options {
	unsigned crop_overscan_pixels;
	unsigned crop_left, crop_right, crop_top, crop_bottom;
	int crop_enable;
	int persist_frame_on_invalid_input;
	int blend_previous_frame;
	int upscale_deinterlacing;
	int vi_aa, vi_scale, vi_serrate, vi_dither_filter, vi_divot_filter, vi_gamma_dither;
} rdp_mi_scanout_options;

void rdp_mi_scanout_options_default(rdp_mi_scanout_options *opts);

/* CommandProcessor::scanout_sync(colors, width, height, opts),
 * rdp_device.cpp:1159-1182, with explicit options. */
int rdp_mi_scanout_sync_ex(rdp_mi_processor *p, const rdp_mi_scanout_options *opts, uint32_t *out, unsigned *width,
                           unsigned *height);

/* VIScanoutBuffer (rdp_device.hpp:53-58): a host-readable image of one
 * scanout with a fence.  One buffer holds up to 640 x 576 RGBA8 words, and
 * grows on first use for an upscaling processor's larger image. */
typedef struct rdp_mi_scanout_buffer rdp_mi_scanout_buffer;
int rdp_mi_scanout_buffer_create(rdp_mi_scanout_buffer **out);
void rdp_mi_scanout_buffer_destroy(rdp_mi_scanout_buffer *b);

/* CommandProcessor::scanout_async_buffer(VIScanoutBuffer &, opts),
 * rdp_device.cpp:1130-1157: drains the command ring, submits the pending
 * rendering and queues the VI scanout and its copy into the buffer behind
 * it, then returns without waiting.  The buffer's width and height are known
 * at once (0 x 0 for a blank or invalid frame); its pixels once
 * rdp_mi_scanout_buffer_wait returns. */
int rdp_mi_scanout_async_buffer(rdp_mi_processor *p, const rdp_mi_scanout_options *opts, rdp_mi_scanout_buffer *b);

/* Waits for the buffer's fence (VIScanoutBuffer::fence->wait(),
 * rdp_device.cpp:1176) and returns its pixels (valid until the buffer is used
 * again or destroyed) and size. */
int rdp_mi_scanout_buffer_wait(rdp_mi_scanout_buffer *b, const uint32_t **pixels, unsigned *width, unsigned *height);

/* Per-scanline HStart / XScale registers: begin_vi_register_per_scanline,
 * set_vi_register_for_scanline, latch_vi_register_for_scanline,
 * end_vi_register_per_scanline (rdp_device.cpp:977-995,
 * video_interface.cpp:1126-1213).  flags / reg_bit: 1 = HStart, 2 = XScale.
 * Latched lines must increase monotonically; the state applies to the next
 * scanout only. */
void rdp_mi_begin_vi_register_per_scanline(rdp_mi_processor *p, uint32_t flags);
void rdp_mi_set_vi_register_for_scanline(rdp_mi_processor *p, uint32_t reg_bit, uint32_t value);
void rdp_mi_latch_vi_register_for_scanline(rdp_mi_processor *p, unsigned vi_line);
void rdp_mi_end_vi_register_per_scanline(rdp_mi_processor *p);

/* ValidationInterface (rdp_renderer.hpp:75-95), installed with
 * CommandProcessor::set_validation_interface (rdp_device.cpp:230-233).
 * Reports RDP states that would hang real hardware; never aborts. */
typedef void (*rdp_mi_validation_fn)(uint32_t error, const char *message, void *user);
void rdp_mi_set_validation_callback(rdp_mi_processor *p, rdp_mi_validation_fn fn, void *user);

/* Batched submission across processors (no reference counterpart; the
 * reference runs one processor per device queue).  Submits the pending
 * render passes of all processors with one launch per stage per pass round,
 * so independent streams share every kernel launch.  Results are identical
 * to flushing each processor on its own. */
int rdp_mi_batch_flush(rdp_mi_processor *const *procs, unsigned count);

/* Allocation-free rdp_mi_batch_flush.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * launches as rdp_mi_batch_flush, with the packed inputs and per-pass
 * scratch in a static device array of the library (16 MiB of packed inputs
 * plus 112 MiB of scratch; each GPU the library runs on holds it until the
 * process ends) and the host side of the upload in a static host array: the
 * call allocates no device or pinned host memory.  It returns when the
 * device has finished the submitted work.  A submission beyond either limit
 * returns RDP_MI_ERROR_LIMIT before any copy or launch and leaves the passes
 * pending (rdp_mi_batch_flush or rdp_mi_flush then submits them).  Every
 * call on one GPU shares the array; calls are serialised by the library's
 * context lock, and the kernels write every scratch byte before reading it,
 * so nothing needs clearing between calls. */
int rdp_mi_batch_flush_noalloc(rdp_mi_processor *const *procs, unsigned count);

/* Recorded submissions (no reference counterpart; the analogue of a
 * pre-recorded command buffer).  rdp_mi_record takes the pending render
 * passes of the processors (batched != 0: one launch per stage per pass
 * round across all processors; 0: one launch per stage per pass), uploads
 * their inputs once and keeps them device-resident.  A replay re-executes
 * the recorded passes on the processors' current RDRAM and TMEM, either as
 * direct launches or as one captured hipGraph; iteration_ms (optional)
 * receives each replay's GPU time.  rdp_mi_recording_profile times the four
 * stages (setup, TMEM, bin, raster) with events between direct launches.
 * The profile and autotune calls execute the recording many times; they copy
 * the recorded processors' RDRAM, hidden RDRAM and TMEM aside on the device
 * first and put them back afterwards, so they leave the state as they found
 * it (a replay changes it).  The processors must outlive the recording. */
typedef struct rdp_mi_recording rdp_mi_recording;
int rdp_mi_record(rdp_mi_processor *const *procs, unsigned count, int batched, rdp_mi_recording **out);
int rdp_mi_recording_replay(rdp_mi_recording *r, unsigned iterations, int use_graph, float *iteration_ms);
int rdp_mi_recording_profile(rdp_mi_recording *r, unsigned iterations, float stage_ms[4]);
/* Per-round variant: round_stage_ms holds rounds * 4 entries. */
int rdp_mi_recording_profile_rounds(rdp_mi_recording *r, unsigned iterations, float *round_stage_ms);
/* Every raster launch of the recording timed alone (stages in series, events
 * around each launch), averaged over `iterations` replays.  For launch k, in
 * launch order: ms[k], its algorithmic framebuffer bytes[k] (colour, depth
 * and hidden bits read and written once per tile), kind[k] (0-2: the
 * k_raster occupancy class, 3: k_raster_fill) and tiles[k].  *count receives
 * the number of raster launches; at most max entries are written. */
int rdp_mi_recording_profile_raster(rdp_mi_recording *r, unsigned iterations, unsigned max, float *ms,
                                    uint64_t *bytes, uint32_t *kind, uint32_t *tiles, unsigned *count);
/* Launch-shape autotuner (no reference counterpart; the reference picks its
 * shader variants and workgroup layout statically, rdp_renderer.cpp:1790-1822).
 * Times the recorded submission under every candidate raster launch shape:
 * the XCD chunk of the tile-to-workgroup mapping (-1 block order, 0 one
 * contiguous eighth per XCD, 32/64/128 tiles per chunk), the lowest
 * occupancy class passes run in (0-2), the tile a workgroup takes
 * (tile_size 8: one 8x8 tile, one wave; 16: a 16x16 tile, one wave per 8x8
 * quadrant; the reference's binner anticipates both block sizes,
 * tile_binning_combined.comp:23) and whether a round's raster work runs as
 * one persistent launch (persistent = 1: resident one-wave workgroups stride
 * over the round's 8x8 tiles, heaviest class first; rounds with top-class
 * passes keep the per-class launches); keeps the fastest for later
 * replays and reports it.  Each candidate is timed over `iterations`
 * replays, best of two sweeps. */
int rdp_mi_recording_autotune(rdp_mi_recording *r, unsigned iterations, int32_t *xcd_chunk, uint32_t *min_class,
                              uint32_t *tile_size, int32_t *persistent, float *best_ms);
/* The last autotune's candidates and their milliseconds per replay (no
 * reference counterpart): writes up to max entries, returns how many
 * candidates there are (0 before an autotune). */
unsigned rdp_mi_recording_autotune_results(const rdp_mi_recording *r, unsigned max, int32_t *xcd_chunk,
                                           uint32_t *min_class, uint32_t *tile_size, int32_t *persistent, float *ms);
/* Launch shape for immediate submissions (flush, batch flush, serving), for
 * example the one an autotuned recording of the same workload picked.
 * tile_size is 8 or 16, and persistent needs 8 (RDP_MI_ERROR_INVALID_ARGUMENT
 * otherwise).  Default: xcd_chunk 64, min_class 0, tile_size 8, persistent 0. */
int rdp_mi_set_launch_config(int32_t xcd_chunk, uint32_t min_class, uint32_t tile_size, int persistent);
uint64_t rdp_mi_recording_raster_bytes(const rdp_mi_recording *r);
unsigned rdp_mi_recording_rounds(const rdp_mi_recording *r);
unsigned rdp_mi_recording_passes(const rdp_mi_recording *r);
/* Raster kernel launches per replay (one per occupancy class present in a
 * round, after the launch shape's class floor). */
unsigned rdp_mi_recording_raster_launches(const rdp_mi_recording *r);
void rdp_mi_recording_destroy(rdp_mi_recording *r);

/* Serving (no reference counterpart: the reference serves one emulator per
 * device queue).  A server takes over the scheduling of `count` processors
 * and renders frame requests with continuous batching:
 *  - a request that has arrived is admitted as soon as its processor is idle
 *    and the estimated device bytes of admitted, not yet submitted frames
 *    stay within hbm_budget_bytes (0 = 4 GiB); it is decoded at once on one
 *    of decode_threads host threads (0 = a quarter of the host threads, at most 8);
 *  - decoded frames are submitted together (at most max_batch per batch,
 *    0 = count, within the budget) whenever one of `streams` submission
 *    streams (0 = 3, at most 8) is free, so new work starts while earlier
 *    batches still run; each stream keeps its own input/scratch arena.
 * Each processor renders its requests in arrival order, one at a time, and
 * its results are identical to rendering the frames alone.  While a server
 * exists, its processors must not be driven through the other entry
 * points. */
typedef struct rdp_mi_server rdp_mi_server;
int rdp_mi_server_create(rdp_mi_processor *const *procs, unsigned count, size_t hbm_budget_bytes, unsigned max_batch,
                         unsigned decode_threads, unsigned streams, rdp_mi_server **out);
/* Queue one frame (complete commands) for processor `instance`, arriving
 * arrival_us microseconds after rdp_mi_server_start (requests queued before
 * the start keep their schedule; after the start, earlier times mean now). */
int rdp_mi_server_submit(rdp_mi_server *s, unsigned instance, const uint32_t *words, size_t num_words,
                         double arrival_us, uint64_t *ticket);
int rdp_mi_server_start(rdp_mi_server *s);
/* Block until every submitted request has completed. */
int rdp_mi_server_drain(rdp_mi_server *s);
/* After a drain: forget completed requests and counters and start a new
 * time base start_in_us microseconds from now (arrival times of later
 * requests count from it).  The server keeps its threads, streams and
 * arenas, so measurements after a reset see it warm. */
int rdp_mi_server_reset(rdp_mi_server *s, double start_in_us);
/* Completed requests in completion order: ticket, processor, arrival,
 * admission, submission (decoded and queued to the GPU) and completion times
 * (microseconds since the start).  Returns the number completed so far;
 * copies at most max.  Any output pointer may be null. */
size_t rdp_mi_server_results(rdp_mi_server *s, uint64_t *tickets, uint32_t *instances, double *arrival_us,
                             double *admit_us, double *submit_us, double *complete_us, size_t max);
/* Batching window (no reference counterpart): a free submission stream
 * takes the decoded frames only once max_batch of them are ready, no
 * admitted frame is still decoding, or the oldest ready frame has waited
 * window_us microseconds.  0 (the default) submits whatever is ready as soon
 * as a stream is free.  Trades latency for larger batches. */
int rdp_mi_server_set_batch_window(rdp_mi_server *s, double window_us);
/* Per-frame state digests (no reference counterpart; the reference's
 * conformance driver compares all of RDRAM after each test,
 * conformance_utils.hpp:282-304).  Enabled before rdp_mi_server_start: after
 * every batch the server computes, on the device and behind the batch's
 * launches, the digest of each served processor's RDRAM, hidden RDRAM and
 * TMEM (the definition of rdp_mi_state_digest), so each frame can be checked
 * on its own. */
int rdp_mi_server_set_digests(rdp_mi_server *s, int enable);
/* Completed requests in the order of rdp_mi_server_results: ticket, the
 * digest after the frame (0 when digests are off) and a status (0, or the
 * negative RDP_MI_ERROR_* code of a batch that failed). */
size_t rdp_mi_server_digests(rdp_mi_server *s, uint64_t *tickets, uint64_t *digests, int32_t *status, size_t max);

/* Digest of a processor's state (no reference counterpart), computed on the
 * device after the work already submitted for it: the sum modulo 2^64 over
 * the 32-bit words w_i of RDRAM (host order), of hidden RDRAM (bytes packed
 * four to a little-endian word) and of TMEM (halfwords packed two to a
 * little-endian word) of SplitMix64(key_r ^ (i << 32 | w_i)), i counting
 * from 0 in each region, key_r = 0x243f6a8885a308d3 (RDRAM),
 * 0x13198a2e03707344 (hidden), 0xa4093822299f31d0 (TMEM).  Call it after
 * rdp_mi_idle for the state the reference's readback would show. */
int rdp_mi_state_digest(rdp_mi_processor *p, uint64_t *digest);

/* Counters so far: out[0] batches submitted, out[1] frames, out[2] frames in
 * the largest batch, out[3] largest estimated batch bytes (the admission
 * figure), out[4] largest actual packed-input plus scratch bytes of a batch,
 * out[5] total and out[6] longest host time spent submitting a batch (us). */
void rdp_mi_server_stats(rdp_mi_server *s, uint64_t out[7]);
void rdp_mi_server_destroy(rdp_mi_server *s);

/* Framebuffer views for capture inspection (the colour / depth / coverage
 * views of the reference's capture replayer, rdp_replayer.cpp:246-393):
 * width x height RGBA8 pixels (R in the low byte) decoded from the
 * processor's RDRAM and hidden RDRAM after the work enqueued so far.
 *  - RDP_MI_VIEW_COLOR: the image at fb_addr; fb_size is SetColorImage's
 *    size field (2: RGBA5551 with 5-bit channels replicated to 8 and alpha
 *    from bit 0; 3: RGBA8888; 1: 8-bit as grey, which the reference omits).
 *  - RDP_MI_VIEW_DEPTH: the 16-bit words at depth_addr with the two low bits
 *    cleared, shown as RGBA5551, and the 4-bit dz (two hidden bits below the
 *    word's two low bits) in alpha.
 *  - RDP_MI_VIEW_COVERAGE: for 16-bit colour images, the 3-bit coverage
 *    (bit 0 over the two hidden bits) as grey; black otherwise.
 * height 0 takes width * 3 / 4, as the reference assumes 4:3. */
enum { RDP_MI_VIEW_COLOR = 0, RDP_MI_VIEW_DEPTH = 1, RDP_MI_VIEW_COVERAGE = 2 };
int rdp_mi_framebuffer_view(rdp_mi_processor *p, uint32_t mode, uint32_t fb_addr, uint32_t fb_size,
                            uint32_t depth_addr, uint32_t width, uint32_t height, uint32_t *out);

/* ---- RDPDUMP2 captures ----
 * The reference's capture format (writer rdp_dump_write.hpp:31-57 /
 * rdp_dump_write.cpp:37-145, reader rdp_dump.hpp:49-68 / rdp_dump.cpp:43-200),
 * byte for byte: "RDPDUMP2", u32 RDRAM size, u32 hidden RDRAM size, then
 * records of host-order u32 words (kinds below).
 *
 * Capture from a processor: rdp_mi_create starts one when the environment
 * variable PARALLEL_RDP_DUMP_PATH names a file (rdp_device.cpp:60-73), or
 * rdp_mi_begin_dump starts one later.  Then, as in rdp_device.cpp:874-897,
 * 970-975 and 1087-1092: the first rdp_mi_enqueue_command of a command list
 * waits for the processor to go idle and records the 4 KiB blocks of RDRAM
 * and hidden RDRAM that changed since the last record; each command is
 * recorded, except SyncFull, which is recorded as SignalComplete and ends the
 * list; VI register writes are recorded; a scanout records the changed
 * blocks and EndFrame.  rdp_mi_enqueue_command_direct is not captured (as in
 * the reference).  rdp_mi_end_dump (or rdp_mi_destroy) writes EndOfFile. */
enum {
	RDP_MI_DUMP_EOF = 0,  /* reported by rdp_mi_dump_player_next at the end */
	RDP_MI_DUMP_UPDATE_DRAM = 1,
	RDP_MI_DUMP_RDP_COMMAND = 2,
	RDP_MI_DUMP_SET_VI_REGISTER = 3,
	RDP_MI_DUMP_END_FRAME = 4,
	RDP_MI_DUMP_SIGNAL_COMPLETE = 5,
	RDP_MI_DUMP_END_OF_FILE = 6,
	RDP_MI_DUMP_UPDATE_DRAM_FLUSH = 7,
	RDP_MI_DUMP_UPDATE_HIDDEN_DRAM = 8,
	RDP_MI_DUMP_UPDATE_HIDDEN_DRAM_FLUSH = 9
};
int rdp_mi_begin_dump(rdp_mi_processor *p, const char *path);
void rdp_mi_end_dump(rdp_mi_processor *p);

/* RDPDumpWriter (rdp_dump_write.hpp:31-57) on its own, for front ends that
 * write captures themselves. */
typedef struct rdp_mi_dump_writer rdp_mi_dump_writer;
int rdp_mi_dump_writer_open(const char *path, uint32_t rdram_size, uint32_t hidden_rdram_size, rdp_mi_dump_writer **out);
void rdp_mi_dump_writer_flush_rdram(rdp_mi_dump_writer *w, const void *rdram, uint32_t size);
void rdp_mi_dump_writer_flush_hidden_rdram(rdp_mi_dump_writer *w, const void *hidden, uint32_t size);
void rdp_mi_dump_writer_command(rdp_mi_dump_writer *w, uint32_t op, const uint32_t *words, uint32_t num_words);
void rdp_mi_dump_writer_set_vi_register(rdp_mi_dump_writer *w, uint32_t reg, uint32_t value);
void rdp_mi_dump_writer_end_frame(rdp_mi_dump_writer *w);
void rdp_mi_dump_writer_signal_complete(rdp_mi_dump_writer *w);
/* Writes EndOfFile, closes the file and frees the writer. */
void rdp_mi_dump_writer_close(rdp_mi_dump_writer *w);

/* DumpPlayer (rdp_dump.hpp:49-68): rdp_mi_dump_player_next reads one record
 * (returns 1; 0 with kind RDP_MI_DUMP_EOF at the end; a negative error for a
 * malformed record).  Block records update the player's RDRAM and hidden
 * RDRAM shadows (rdp_mi_dump_player_rdram / _hidden_rdram); `words` stays
 * valid until the next call.  rdp_mi_dump_player_apply drives a processor
 * with one record as the reference's replay driver does
 * (replayer_driver_parallel.cpp:124-205): commands are enqueued, VI
 * registers set, SignalComplete flushes, and each UpdateDramFlush /
 * UpdateHiddenDramFlush waits for idle and copies the shadow into the
 * processor's RDRAM / hidden RDRAM; EndFrame is left to the caller (the
 * reference's driver scans out there with blend_previous_frame = 1 and
 * upscale_deinterlacing = 0). */
typedef struct rdp_mi_dump_player rdp_mi_dump_player;
typedef struct rdp_mi_dump_event {
	uint32_t kind;
	uint32_t op, num_words;
	const uint32_t *words;
	uint32_t vi_register, vi_value;
	uint32_t offset, size;
} rdp_mi_dump_event;
int rdp_mi_dump_player_open(const char *path, rdp_mi_dump_player **out);
uint32_t rdp_mi_dump_player_rdram_size(const rdp_mi_dump_player *pl);
uint32_t rdp_mi_dump_player_hidden_rdram_size(const rdp_mi_dump_player *pl);
int rdp_mi_dump_player_rewind(rdp_mi_dump_player *pl);
int rdp_mi_dump_player_next(rdp_mi_dump_player *pl, rdp_mi_dump_event *ev);
const uint8_t *rdp_mi_dump_player_rdram(const rdp_mi_dump_player *pl);
const uint8_t *rdp_mi_dump_player_hidden_rdram(const rdp_mi_dump_player *pl);
int rdp_mi_dump_player_apply(rdp_mi_dump_player *pl, const rdp_mi_dump_event *ev, rdp_mi_processor *p);
void rdp_mi_dump_player_destroy(rdp_mi_dump_player *pl);

/* Library diagnostics. */
const char *rdp_mi_last_error(void);
uint64_t rdp_mi_primitive_count(const rdp_mi_processor *p);

#ifdef __cplusplus
}
#endif

#endif

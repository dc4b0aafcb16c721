// Test-only C shim (no device code): runs the host front end over command
// words and checks, for every TMEM upload of every pass, that each RDRAM byte
// the upload plan reads lies inside upload_read_range, the bound the
// submission uses to let a round's TMEM replay start before the previous
// round's raster.  Built by tests/test_hazard.py with g++.
#include <stdint.h>
#include <stddef.h>

#include "../parallel-rdp_amd/csrc/rdp_frontend.hpp"
#include "../parallel-rdp_amd/csrc/rdp_hazard.hpp"
#include "../parallel-rdp_amd/csrc/rdp_tmem_plan.hpp"

using namespace rdp_mi;

extern "C" {
// out[0] = uploads, out[1] = bytes read, out[2] = bytes outside the bound,
// out[3] = uploads the bound refuses (no bound: never start early).
int hz_check(const uint32_t *words, size_t n, uint32_t rdram_size, uint64_t out[4])
{
	Frontend fe(rdram_size);
	size_t i = 0;
	while (i < n) {
		unsigned len = Frontend::command_length(words[i]);
		if (len == 0) { i++; continue; }
		if (i + len > n) break;
		fe.command(words + i);
		i += len;
	}
	fe.flush();
	out[0] = out[1] = out[2] = out[3] = 0;
	for (const HostPass &hp : fe.completed()) {
		for (const UploadInfo &u : hp.uploads) {
			out[0]++;
			ByteRange b;
			if (!upload_read_range(u, b)) {
				out[3]++;
				continue;
			}
			const uint64_t lo = b.lo % rdram_size, hi = lo + (b.hi - b.lo);
			for (int g = 0; g < 2048; g++) {
				// The gather's lane-to-halfword order (tmem_update.comp).
				const int idx = g ^ 1;
				TmemLoad ld;
				if (!tmem_plan(u, idx, idx >= 0x400, ld)) continue;
				const int nb = ld.kind == TMEM_LOAD_16 ? 2 : 4;
				for (int q = 0; q < nb; q++) {
					const uint64_t a = (uint64_t)(ld.addr + q) & (rdram_size - 1);
					out[1]++;
					const bool inside = (a >= lo && a < hi) || (hi > rdram_size && a < hi - rdram_size);
					if (!inside && b.hi - b.lo < rdram_size) out[2]++;
				}
			}
		}
	}
	return 0;
}
}

// The other half of the bound: every RDRAM byte (and hidden byte) a device
// pass's raster changes lies inside pass_write_ranges of that pass.  The
// oracle renders the stream; its state is taken right before the first draw
// of each device pass (device passes end where the reference's renderer ends
// a render pass, so flushing the oracle there changes nothing) and at the end.
#include "../oracle/rdp_oracle.h"
#include <string.h>

static bool inside_ranges(const std::vector<ByteRange> &rs, uint64_t a, uint32_t size)
{
	for (const ByteRange &r : rs) {
		const uint64_t lo = r.lo % size, hi = lo + (r.hi - r.lo);
		if (r.hi - r.lo >= size) return true;
		if ((a >= lo && a < hi) || (hi > size && a < hi - size)) return true;
	}
	return false;
}

extern "C" {
// out[0] = device passes with draws, out[1] = changed bytes seen,
// out[2] = changed bytes outside the pass's write ranges.
int hz_write_check(const uint32_t *words, size_t n, uint32_t rdram_size, const uint32_t *rdram_init,
                   uint32_t hidden_init, uint64_t out[3])
{
	std::vector<HostPass> passes;
	{
		Frontend fe(rdram_size);
		size_t i = 0;
		while (i < n) {
			unsigned len = Frontend::command_length(words[i]);
			if (len == 0) { i++; continue; }
			if (i + len > n) break;
			fe.command(words + i);
			i += len;
		}
		fe.flush();
		for (HostPass &hp : fe.completed())
			if (!hp.prims.empty()) passes.push_back(std::move(hp));
	}
	out[0] = passes.size();
	out[1] = out[2] = 0;
	rdpo *o = rdpo_create(rdram_size);
	if (rdram_init) memcpy(rdpo_rdram(o), rdram_init, rdram_size);
	memset(rdpo_hidden_rdram(o), (int)hidden_init, rdram_size / 2);
	std::vector<uint8_t> prev(rdpo_rdram(o), rdpo_rdram(o) + rdram_size);
	std::vector<uint8_t> prev_h(rdpo_hidden_rdram(o), rdpo_hidden_rdram(o) + rdram_size / 2);
	size_t next = 0;  // index of the next device pass whose start is pending
	auto close_pass = [&](size_t k) {
		// Bytes changed since the start of pass k belong to pass k.
		rdpo_flush(o);
		const uint8_t *cur = rdpo_rdram(o), *cur_h = rdpo_hidden_rdram(o);
		std::vector<ByteRange> rs;
		pass_write_ranges(passes[k], rs);
		for (uint32_t b = 0; b < rdram_size; b++)
			if (cur[b] != prev[b]) {
				out[1]++;
				// Host byte b holds RDRAM byte b ^ 3.
				if (!inside_ranges(rs, b ^ 3u, rdram_size)) out[2]++;
			}
		for (uint32_t h = 0; h < rdram_size / 2; h++)
			if (cur_h[h] != prev_h[h]) {
				out[1]++;
				if (!inside_ranges(rs, 2ull * h, rdram_size)) out[2]++;
			}
		memcpy(prev.data(), cur, rdram_size);
		memcpy(prev_h.data(), cur_h, rdram_size / 2);
	};
	size_t i = 0;
	while (i < n) {
		unsigned len = Frontend::command_length(words[i]);
		if (len == 0) { i++; continue; }
		if (i + len > n) break;
		const uint64_t before = rdpo_primitive_count(o);
		const unsigned op = (words[i] >> 24) & 63;
		const bool draw = (op >= 0x08 && op <= 0x0f) || op == 0x24 || op == 0x25 || op == 0x36;
		if (draw && next < passes.size() && before == passes[next].prims.front().prim_index) {
			// The next command is the first draw of device pass `next`:
			// everything so far belongs to the passes before it.
			if (next > 0) close_pass(next - 1);
			else {
				rdpo_flush(o);
				if (memcmp(prev.data(), rdpo_rdram(o), rdram_size) != 0) out[2]++;
			}
			next++;
		}
		rdpo_process(o, words + i, len);
		i += len;
	}
	if (next > 0) close_pass(next - 1);
	if (next != passes.size()) out[2] += 1000000;  // a pass start never seen
	rdpo_destroy(o);
	return 0;
}
}

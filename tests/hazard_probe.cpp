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

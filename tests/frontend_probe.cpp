// Test-only C shim over the host front end (no device code): feeds command
// words and reports the passes it builds and the reference pass boundaries
// it tracks.  Built by tests/test_frontend.py with g++.
#include <stdint.h>
#include <stddef.h>

#include <vector>

#include "../parallel-rdp_amd/csrc/rdp_frontend.hpp"

using rdp_mi::Frontend;

struct Probe {
	Frontend fe;
	std::vector<uint64_t> log;
	explicit Probe(uint32_t size) : fe(size) { fe.pass_end_log = &log; }
};

extern "C" {
void *fp_create(uint32_t rdram_size) { return new Probe(rdram_size); }
void fp_destroy(void *p) { delete static_cast<Probe *>(p); }

void fp_process(void *h, const uint32_t *words, size_t n)
{
	Probe *p = static_cast<Probe *>(h);
	size_t i = 0;
	while (i < n) {
		unsigned len = Frontend::command_length(words[i]);
		if (len == 0) { i++; continue; }
		if (i + len > n) break;
		p->fe.command(words + i);
		i += len;
	}
}

void fp_flush(void *h) { static_cast<Probe *>(h)->fe.flush(); }

size_t fp_pass_ends(void *h, uint64_t *out, size_t cap)
{
	Probe *p = static_cast<Probe *>(h);
	for (size_t i = 0; i < p->log.size() && i < cap; i++) out[i] = p->log[i];
	return p->log.size();
}

// Per completed pass: prim_index of its first and last primitive record
// (0xffffffff when it holds only uploads) and its upload count.
size_t fp_passes(void *h, uint32_t *first, uint32_t *last, uint32_t *uploads, size_t cap)
{
	Probe *p = static_cast<Probe *>(h);
	auto &done = p->fe.completed();
	for (size_t i = 0; i < done.size() && i < cap; i++) {
		first[i] = done[i].prims.empty() ? 0xffffffffu : done[i].prims.front().prim_index;
		last[i] = done[i].prims.empty() ? 0xffffffffu : done[i].prims.back().prim_index;
		uploads[i] = (uint32_t)done[i].uploads.size();
	}
	return done.size();
}
}

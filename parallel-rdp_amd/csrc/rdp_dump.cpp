// RDPDUMP2 capture writer and reader (format: rdp_dump.hpp).
#include "rdp_dump.hpp"

#include <string.h>

#include <algorithm>

namespace rdp_mi {

static constexpr uint32_t kDumpBlock = 4096;

bool DumpWriter::open(const char *path, uint32_t rdram_size, uint32_t hidden_size)
{
	if (file || !path) return false;
	file = fopen(path, "wb");
	if (!file) return false;
	// The shadows start zeroed, as the reader's do, so the first flush sends
	// every non-zero block.
	shadow[0].assign(rdram_size, 0);
	shadow[1].assign(hidden_size, 0);
	fwrite("RDPDUMP2", 1, 8, file);
	put(rdram_size);
	put(hidden_size);
	return true;
}

void DumpWriter::close()
{
	if (!file) return;
	put(DUMP_EOF);
	fclose(file);
	file = nullptr;
	shadow[0].clear();
	shadow[1].clear();
}

void DumpWriter::memory(const uint8_t *data, uint32_t size, bool hidden)
{
	if (!file) return;
	std::vector<uint8_t> &sh = shadow[hidden ? 1 : 0];
	size = std::min<uint32_t>(size, (uint32_t)sh.size());
	for (uint32_t off = 0; off < size; off += kDumpBlock) {
		const uint32_t n = std::min(kDumpBlock, size - off);
		if (memcmp(data + off, sh.data() + off, n) == 0) continue;
		put(hidden ? DUMP_UPDATE_HIDDEN_DRAM : DUMP_UPDATE_DRAM);
		put(off);
		put(n);
		fwrite(data + off, 1, n, file);
		memcpy(sh.data() + off, data + off, n);
	}
	put(hidden ? DUMP_UPDATE_HIDDEN_DRAM_FLUSH : DUMP_UPDATE_DRAM_FLUSH);
}

void DumpWriter::command(uint32_t op, const uint32_t *words, uint32_t num_words)
{
	if (!file) return;
	put(DUMP_RDP_COMMAND);
	put(op);
	put(num_words);
	fwrite(words, sizeof(uint32_t), num_words, file);
}

void DumpWriter::vi_register(uint32_t reg, uint32_t value)
{
	if (!file) return;
	put(DUMP_SET_VI_REGISTER);
	put(reg);
	put(value);
}

DumpReader::~DumpReader()
{
	if (file) fclose(file);
}

bool DumpReader::open(const char *path)
{
	if (file || !path) return false;
	file = fopen(path, "rb");
	if (!file) return false;
	char magic[8];
	uint32_t rdram_size = 0, hidden_size = 0;
	if (fread(magic, 1, 8, file) != 8 || memcmp(magic, "RDPDUMP2", 8) != 0 || !get(rdram_size) || !get(hidden_size) ||
	    (rdram_size != (4u << 20) && rdram_size != (8u << 20)) ||
	    (hidden_size != (4u << 20) && hidden_size != rdram_size / 2)) {
		fclose(file);
		file = nullptr;
		return false;
	}
	shadow[0].assign(rdram_size, 0);
	shadow[1].assign(hidden_size, 0);
	return true;
}

bool DumpReader::rewind()
{
	if (!file || fseek(file, 16, SEEK_SET) != 0) return false;
	std::fill(shadow[0].begin(), shadow[0].end(), 0);
	std::fill(shadow[1].begin(), shadow[1].end(), 0);
	return true;
}

int DumpReader::next(DumpEvent &ev)
{
	if (!file) return -1;
	ev = DumpEvent();
	uint32_t kind;
	if (!get(kind)) return 0;
	ev.kind = kind;
	switch (kind) {
	case DUMP_EOF: return 0;
	case DUMP_END_FRAME:
	case DUMP_SIGNAL_COMPLETE:
	case DUMP_UPDATE_DRAM_FLUSH:
	case DUMP_UPDATE_HIDDEN_DRAM_FLUSH: return 1;
	case DUMP_SET_VI_REGISTER: return get(ev.vi_register) && get(ev.vi_value) ? 1 : -1;
	case DUMP_RDP_COMMAND: {
		uint32_t n;
		if (!get(ev.op) || !get(n) || n > 1024) return -1;
		words.resize(std::max<uint32_t>(n, 1));
		if (n && fread(words.data(), sizeof(uint32_t), n, file) != n) return -1;
		ev.num_words = n;
		ev.words = words.data();
		return 1;
	}
	case DUMP_UPDATE_DRAM:
	case DUMP_UPDATE_HIDDEN_DRAM: {
		std::vector<uint8_t> &sh = shadow[kind == DUMP_UPDATE_HIDDEN_DRAM ? 1 : 0];
		if (!get(ev.offset) || !get(ev.size)) return -1;
		if ((uint64_t)ev.offset + ev.size > sh.size()) return -1;
		if (ev.size && fread(sh.data() + ev.offset, 1, ev.size, file) != ev.size) return -1;
		return 1;
	}
	default: return -1;
	}
}

}  // namespace rdp_mi

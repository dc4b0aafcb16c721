// RDPDUMP2 command captures: writer and reader.
//
// The file format is the reference's (rdp_dump_write.cpp:37-145 writes it,
// rdp_dump.cpp:43-200 reads it), kept byte for byte so captures move between
// the two implementations.  Every field is a host-order u32:
//
//   "RDPDUMP2"  rdram_size  hidden_rdram_size
//   then records, each starting with its kind:
//     1 UpdateDram            offset size bytes[size]   (RDRAM shadow block)
//     2 RDPCommand            op num_words words[num_words]
//     3 SetVIRegister         register value
//     4 EndFrame
//     5 SignalComplete        (a SyncFull: the command list ends here)
//     6 EndOfFile
//     7 UpdateDramFlush       (the RDRAM shadow is now the processor's RDRAM)
//     8 UpdateHiddenDram      offset size bytes[size]
//     9 UpdateHiddenDramFlush
//
// The writer sends only the 4 KiB blocks of RDRAM (and hidden RDRAM) that
// changed since its last flush, as the reference does.
#pragma once

#include <stdint.h>
#include <stdio.h>

#include <vector>

namespace rdp_mi {

enum DumpRecord : uint32_t {
	DUMP_UPDATE_DRAM = 1,
	DUMP_RDP_COMMAND = 2,
	DUMP_SET_VI_REGISTER = 3,
	DUMP_END_FRAME = 4,
	DUMP_SIGNAL_COMPLETE = 5,
	DUMP_EOF = 6,
	DUMP_UPDATE_DRAM_FLUSH = 7,
	DUMP_UPDATE_HIDDEN_DRAM = 8,
	DUMP_UPDATE_HIDDEN_DRAM_FLUSH = 9,
};

class DumpWriter {
public:
	~DumpWriter() { close(); }
	bool open(const char *path, uint32_t rdram_size, uint32_t hidden_size);
	// Writes the EndOfFile record and closes the file.
	void close();
	bool is_open() const { return file != nullptr; }

	// Changed 4 KiB blocks of `data` against the shadow, then the flush
	// record.  hidden selects hidden RDRAM.
	void memory(const uint8_t *data, uint32_t size, bool hidden);
	void command(uint32_t op, const uint32_t *words, uint32_t num_words);
	void vi_register(uint32_t reg, uint32_t value);
	void end_frame() { put(DUMP_END_FRAME); }
	void signal_complete() { put(DUMP_SIGNAL_COMPLETE); }

private:
	FILE *file = nullptr;
	std::vector<uint8_t> shadow[2];
	void put(uint32_t w) { fwrite(&w, sizeof(w), 1, file); }
};

// One record of a capture.  Pointers stay valid until the next call.
struct DumpEvent {
	uint32_t kind = 0;
	uint32_t op = 0, num_words = 0;
	const uint32_t *words = nullptr;
	uint32_t vi_register = 0, vi_value = 0;
	uint32_t offset = 0, size = 0;  // UpdateDram / UpdateHiddenDram
};

class DumpReader {
public:
	~DumpReader();
	// Checks the header: RDRAM 4 or 8 MiB, hidden RDRAM 4 MiB or half the
	// RDRAM.
	bool open(const char *path);
	// Back to the first record, with both shadows cleared.
	bool rewind();
	// 1: an event; 0: end of the capture (EndOfFile, or the file ends on a
	// record boundary); -1: a malformed or truncated record.
	int next(DumpEvent &ev);
	uint32_t rdram_size() const { return (uint32_t)shadow[0].size(); }
	uint32_t hidden_size() const { return (uint32_t)shadow[1].size(); }
	// The RDRAM / hidden RDRAM the capture holds after the records so far.
	const uint8_t *rdram() const { return shadow[0].data(); }
	const uint8_t *hidden() const { return shadow[1].data(); }

private:
	FILE *file = nullptr;
	std::vector<uint8_t> shadow[2];
	std::vector<uint32_t> words;
	bool get(uint32_t &w) { return fread(&w, sizeof(w), 1, file) == 1; }
};

}  // namespace rdp_mi

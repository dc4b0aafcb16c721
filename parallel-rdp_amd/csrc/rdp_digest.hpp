// Device-side digest of processor state (see rdp_digest.hip for the
// definition).  Shared by the host runtime and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rdp_mi {

// One processor's state regions.
struct DigestTarget {
	const uint32_t *rdram;
	const uint8_t *hidden;
	const uint16_t *tmem;  // 2048 halfwords
	uint32_t rdram_words;
	uint32_t hidden_bytes;
};

// Keys that keep the three regions' terms apart (RDRAM, hidden RDRAM, TMEM).
constexpr uint64_t kDigestKeyRdram = 0x243f6a8885a308d3ull;
constexpr uint64_t kDigestKeyHidden = 0x13198a2e03707344ull;
constexpr uint64_t kDigestKeyTmem = 0xa4093822299f31d0ull;

// Zeroes d_out[0 .. n) and adds each target's digest into its slot; max_words
// is the largest rdram_words + hidden_bytes / 4 + 1024 of the targets.
hipError_t launch_digest(const DigestTarget *d_targets, uint32_t n, uint32_t max_words, uint64_t *d_out, hipStream_t s);

}  // namespace rdp_mi

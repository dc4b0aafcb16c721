// Order-independent digest of a processor's device state (RDRAM, hidden
// RDRAM and TMEM), so served frames can be checked one by one without
// reading 12 MiB per frame back to the host.
//
// digest = sum over the state's 32-bit words w_i (mod 2^64) of
//          mix64(region_key[r] ^ (i << 32 | w_i))
// where region_key[r] is kDigestKeyRdram for RDRAM words (host order),
// kDigestKeyHidden for hidden RDRAM bytes packed four to a little-endian word
// and kDigestKeyTmem for TMEM halfwords packed two to a little-endian word, i
// counts words from 0 within the region, and mix64 is the SplitMix64
// finaliser.  A sum of well-mixed terms changes with any
// single word, and blocks can add their partial sums in any order.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rdp_digest.hpp"

namespace rdp_mi {

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
	return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t digest_term(uint64_t key, uint32_t index, uint32_t word)
{
	return mix64(key ^ (((uint64_t)index << 32) | word));
}

constexpr uint32_t kDigestThreads = 256;
constexpr uint32_t kDigestWordsPerThread = 16;
constexpr uint32_t kDigestWordsPerBlock = kDigestThreads * kDigestWordsPerThread;

// Zero the n digest slots.
__global__ __launch_bounds__(64) void k_digest_clear(uint64_t *out, uint32_t n)
{
	for (uint32_t i = threadIdx.x; i < n; i += 64) out[i] = 0;
}

// blockIdx.y: target; blockIdx.x: a run of kDigestWordsPerBlock words of the
// target's concatenated regions.  Consecutive threads read consecutive words.
__global__ __launch_bounds__(kDigestThreads) void k_digest(const DigestTarget *__restrict__ targets,
                                                           uint64_t *__restrict__ out)
{
	const DigestTarget t = targets[blockIdx.y];
	const uint32_t hidden_words = t.hidden_bytes / 4, total = t.rdram_words + hidden_words + 1024;
	const uint32_t base = blockIdx.x * kDigestWordsPerBlock;
	uint64_t acc = 0;
	if (base < total) {
#pragma unroll 4
		for (uint32_t k = 0; k < kDigestWordsPerThread; k++) {
			const uint32_t i = base + k * kDigestThreads + threadIdx.x;
			if (i < t.rdram_words)
				acc += digest_term(kDigestKeyRdram, i, t.rdram[i]);
			else if (i < t.rdram_words + hidden_words) {
				const uint32_t j = i - t.rdram_words;
				acc += digest_term(kDigestKeyHidden, j, reinterpret_cast<const uint32_t *>(t.hidden)[j]);
			} else if (i < total) {
				const uint32_t j = i - t.rdram_words - hidden_words;
				acc += digest_term(kDigestKeyTmem, j, reinterpret_cast<const uint32_t *>(t.tmem)[j]);
			}
		}
	}
	// Wave sum, then one add per wave.
	for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
	if ((threadIdx.x & 63) == 0 && acc) atomicAdd((unsigned long long *)&out[blockIdx.y], (unsigned long long)acc);
}

hipError_t launch_digest(const DigestTarget *d_targets, uint32_t n, uint32_t max_words, uint64_t *d_out, hipStream_t s)
{
	if (n == 0) return hipSuccess;
	hipLaunchKernelGGL(k_digest_clear, dim3(1), dim3(64), 0, s, d_out, n);
	const uint32_t blocks = (max_words + kDigestWordsPerBlock - 1) / kDigestWordsPerBlock;
	hipLaunchKernelGGL(k_digest, dim3(blocks, n), dim3(kDigestThreads), 0, s, d_targets, d_out);
	return hipGetLastError();
}

}  // namespace rdp_mi

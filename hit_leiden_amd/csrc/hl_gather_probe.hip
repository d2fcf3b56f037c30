// hl_gather_probe.hip -- the random-gather ceiling of the device: uniformly
// random 4- and 8-byte loads from tables of 4 MiB to 2 GiB (indices read
// coalesced), in gathers per second.  Every hot kernel of the incremental
// update on power-law graphs is a stream of such gathers (neighbour labels,
// community totals, hub-table slots), so this rate, not the streaming HBM
// bandwidth, is the ceiling they are measured against.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <vector>

#define PROBE_CHECK(x)                                                                   \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) {                                                              \
      std::fprintf(stderr, "%s: %s (line %d)\n", #x, hipGetErrorString(e_), __LINE__);   \
      return 1;                                                                          \
    }                                                                                    \
  } while (0)

template <class T>
__global__ void k_gather(const uint32_t* idx, uint64_t n, const T* table, unsigned long long* sink) {
  unsigned long long acc = 0;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    acc += (unsigned long long)table[idx[i]];
  if (acc == 0x9e3779b97f4a7c15ull) *sink = acc;  // never true: keeps the loads
}

// splitmix64 (a published generator): the host-side index stream
static uint64_t splitmix64(uint64_t& x) {
  uint64_t z = (x += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

int main() {
  const uint64_t n = 1ull << 27;  // gathers per run
  std::vector<uint32_t> h(n);
  uint32_t* didx = nullptr;
  unsigned long long* sink = nullptr;
  PROBE_CHECK(hipMalloc(&didx, n * sizeof(uint32_t)));
  PROBE_CHECK(hipMalloc(&sink, sizeof(unsigned long long)));
  hipEvent_t a, b;
  PROBE_CHECK(hipEventCreate(&a));
  PROBE_CHECK(hipEventCreate(&b));
  std::printf("%-10s %-6s %-14s %-10s\n", "table_MiB", "bytes", "Ggathers/s", "ms");
  for (uint64_t entries : {1ull << 20, 1ull << 22, 1ull << 24, 1ull << 26, 1ull << 28}) {
    uint64_t state = entries;
    for (uint64_t i = 0; i < n; ++i) h[i] = uint32_t(splitmix64(state) % entries);
    PROBE_CHECK(hipMemcpy(didx, h.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
    for (int bytes : {4, 8}) {
      void* t = nullptr;
      PROBE_CHECK(hipMalloc(&t, entries * bytes));
      PROBE_CHECK(hipMemset(t, 1, entries * bytes));
      float ms = 0.f;
      for (int rep = 0; rep < 2; ++rep) {  // the first run warms the TLB and caches
        PROBE_CHECK(hipEventRecord(a));
        if (bytes == 4)
          hipLaunchKernelGGL(k_gather<uint32_t>, 8192, 256, 0, 0, didx, n, static_cast<const uint32_t*>(t), sink);
        else
          hipLaunchKernelGGL(k_gather<uint64_t>, 8192, 256, 0, 0, didx, n, static_cast<const uint64_t*>(t), sink);
        PROBE_CHECK(hipGetLastError());
        PROBE_CHECK(hipEventRecord(b));
        PROBE_CHECK(hipEventSynchronize(b));
        PROBE_CHECK(hipEventElapsedTime(&ms, a, b));
      }
      std::printf("%-10.0f %-6d %-14.2f %-10.3f\n", double(entries * bytes) / 1048576.0, bytes, n / (ms * 1e6), ms);
      PROBE_CHECK(hipFree(t));
    }
  }
  return 0;
}

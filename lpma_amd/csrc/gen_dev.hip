// gen_dev.hip -- R-MAT edge samples generated on the device.
//
// The large workloads (2^28 and 2^29 edges) are drawn on the GPU instead of
// by the host generators: sample i is a pure function of (seed, i), so any
// thread computes any sample and a seed fixes the stream.  The random bits
// come from the SplitMix64 finaliser (Steele, Lea & Flood, "Fast splittable
// pseudorandom number generators", OOPSLA 2014) applied to a counter; each
// 64-bit draw decides two R-MAT levels from its two 32-bit halves.
#include "common.hpp"
#include "kernels.hpp"

namespace pma {

__host__ __device__ __forceinline__ u64 splitmix64(u64 x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// quadrant thresholds on 32-bit draws: u < ta -> (0,0), < tab -> (0,1),
// < tabc -> (1,0), else (1,1)
__global__ void k_gen_rmat(u32 scale, u64 n, u64 ta, u64 tab, u64 tabc, u64 seed, u64 *__restrict__ out) {
    const u64 base = splitmix64(seed);
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) {
        u64 src = 0, dst = 0;
        for (u32 attempt = 0; attempt < 16; ++attempt) {
            u64 stream = splitmix64(base ^ (i * 16 + attempt));
            src = 0;
            dst = 0;
            for (u32 l = 0; l < scale; l += 2) {
                u64 r = splitmix64(stream + l);
                for (u32 h = 0; h < 2 && l + h < scale; ++h) {
                    u64 u = h ? (r >> 32) : (r & 0xFFFFFFFFull);
                    u64 bs = u >= tab ? 1 : 0;
                    u64 bd = (u >= ta && u < tab) || u >= tabc ? 1 : 0;
                    src = (src << 1) | bs;
                    dst = (dst << 1) | bd;
                }
            }
            if (src != dst) break;
        }
        if (src == dst) dst ^= 1;  // 16 self loops in a row: a fixed neighbour instead
        out[i] = (src << 32) | dst;
    }
}

hipError_t gen_rmat_device(u32 scale, u64 n, double a, double b, double c, u64 seed, u64 *out, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const double two32 = 4294967296.0;
    u64 ta = (u64)(a * two32), tab = (u64)((a + b) * two32), tabc = (u64)((a + b + c) * two32);
    k_gen_rmat<<<grid_for(n, 256, 65536), 256, 0, st>>>(scale, n, ta, tab, tabc, seed, out);
    return hipGetLastError();
}

}  // namespace pma

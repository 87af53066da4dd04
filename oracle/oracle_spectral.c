/* FP64 CPU oracle for the pseudo-spectral path -- TEST INFRASTRUCTURE ONLY. */
#include <stdint.h>

/* splitmix64 (src/spectral/ic.jl:31-36) */
uint64_t ospec_splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

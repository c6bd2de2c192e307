// Known-answer vectors for the Philox4x32-10 stream, produced by rocRAND's
// open-source (MIT) cuRAND-compatible device API compiled for the host.
// Build and run with tests/golden/make_philox_kat.sh; output is JSON.
#include <hip/hip_runtime.h>
#include <rocrand/rocrand_kernel.h>
#include <cstdio>

int main() {
    const unsigned long long seeds[] = {0ull, 1ull, 0x123456789abcdefull};
    const unsigned long long subs[] = {0ull, 1ull, 31ull, 32ull, 800001ull, 4294967297ull};
    std::printf("{\n \"raw\": [\n");
    bool first = true;
    for (auto sd : seeds)
        for (auto sb : subs) {
            rocrand_state_philox4x32_10 st;
            rocrand_init(sd, sb, 0, &st);
            std::printf("%s  {\"seed\": %llu, \"sub\": %llu, \"draws\": [", first ? "" : ",\n", sd, sb);
            first = false;
            for (int i = 0; i < 11; ++i) std::printf("%s%u", i ? ", " : "", rocrand(&st));
            std::printf("]}");
        }
    std::printf("\n ],\n \"uniform\": [");
    {
        rocrand_state_philox4x32_10 st;
        rocrand_init(0ull, 33ull, 0, &st);
        for (int i = 0; i < 16; ++i) std::printf("%s%.9g", i ? ", " : "", rocrand_uniform(&st));
    }
    std::printf("],\n \"uniform_seed\": 0, \"uniform_sub\": 33,\n");
    // mixed normal/uniform sequence: exercises the Box-Muller cache.
    const int kinds[] = {1, 1, 1, 0, 0, 0, 1, 1, 1, 0, 1, 0};
    std::printf(" \"mixed_kinds\": [");
    for (int i = 0; i < 12; ++i) std::printf("%s%d", i ? ", " : "", kinds[i]);
    std::printf("],\n \"mixed\": [");
    {
        rocrand_state_philox4x32_10 st;
        rocrand_init(0ull, 65ull, 0, &st);
        for (int i = 0; i < 12; ++i)
            std::printf("%s%.9g", i ? ", " : "", kinds[i] ? rocrand_normal(&st) : rocrand_uniform(&st));
    }
    std::printf("],\n \"mixed_seed\": 0, \"mixed_sub\": 65\n}\n");
    return 0;
}

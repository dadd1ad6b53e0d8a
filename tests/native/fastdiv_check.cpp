// Host check of the device fast division (mg_internal.h) against n / d.
#include "mg_internal.h"

#include <cstdio>

int main() {
    using mg::FastDiv;
    for (uint32_t d = 1; d < 4200; d++) {
        const FastDiv f = mg::make_fastdiv(d);
        for (uint32_t n = 0; n < 200000; n += (n < 50000 ? 1 : 89))
            if (mg::fdiv(n, f) != n / d) { std::printf("FAIL d=%u n=%u\n", d, n); return 1; }
        const uint32_t big[] = {(1u << 31) - 1, (1u << 31) - 2, 123456789u, 2000000000u, (1u << 30) + 7};
        for (uint32_t n : big)
            if (mg::fdiv(n, f) != n / d) { std::printf("FAIL d=%u n=%u\n", d, n); return 1; }
    }
    const uint32_t ds[] = {4096, 32768, 65536, 262144, 1u << 20, 16777216, 1000003, (1u << 31) - 1, 1u << 31,
                           2147483647u / 3, 6 * 6 * 6 * 6 * 6, 3 * 4 * 2 * 7};
    for (uint32_t d : ds) {
        const FastDiv f = mg::make_fastdiv(d);
        for (uint32_t n = 0; n < (1u << 31) - 100000; n += 104729)
            if (mg::fdiv(n, f) != n / d) { std::printf("FAIL d=%u n=%u\n", d, n); return 1; }
        for (uint32_t k = 1; k < 64; k++) {   // around multiples of d
            const uint64_t base = (uint64_t)d * k;
            for (int e = -2; e <= 2; e++) {
                const int64_t n = (int64_t)base + e;
                if (n < 0 || n >= (1ll << 31)) continue;
                if (mg::fdiv((uint32_t)n, f) != (uint32_t)n / d) { std::printf("FAIL d=%u n=%lld\n", d, (long long)n); return 1; }
            }
        }
    }
    std::printf("fastdiv ok\n");
    return 0;
}

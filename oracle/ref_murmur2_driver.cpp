// Known-answer driver for the reference's MurmurHash2 functions (test
// infrastructure).  Linked against the reference's own source,
// unit_tests/bloom_and_hash/sw_murmur2/murmur2.cpp:8-90 (compiled from where
// it lies by oracle/Makefile `ref`, its main() renamed), never copied here.
//
// stdin: one case per line, "<len> <seed> <hex bytes>" (hex may be "-" for
// len 0).  stdout: per line "<MurmurHash2(bytes, len, seed)> <KeyLen4>",
// where KeyLen4 = MurmurHash2_KeyLen4(bytes, seed) for len == 4, else 0.
#include <stdint.h>
#include <stdio.h>
#include <string.h>

uint32_t MurmurHash2(const void *key, int len, uint32_t seed);
uint32_t MurmurHash2_KeyLen4(const void *key, uint32_t seed);

int main()
{
    int len;
    unsigned long seed;
    char hex[256];
    while (scanf("%d %lu %255s", &len, &seed, hex) == 3) {
        unsigned char buf[128];
        uint32_t aligned[32];
        if (len < 0 || len > 100) return 1;
        for (int i = 0; i < len; i++) {
            unsigned v = 0;
            if (sscanf(hex + 2 * i, "%2x", &v) != 1) return 1;
            buf[i] = (unsigned char)v;
        }
        memcpy(aligned, buf, (size_t)len);
        uint32_t h = MurmurHash2(aligned, len, (uint32_t)seed);
        uint32_t h4 = len == 4 ? MurmurHash2_KeyLen4(aligned, (uint32_t)seed) : 0u;
        printf("%u %u\n", h, h4);
    }
    return 0;
}

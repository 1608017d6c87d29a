// Host check of the multiply-shift division used by the lazy gather
// kernels (qob_internal.h): compares fdiv with '/' over divisors 1..4096
// and numerators at and around every multiple boundary plus random draws.
#include "../quantumopticsbase.jl_amd/csrc/qob_internal.h"
#include <cstdio>
#include <random>

int main() {
    std::mt19937_64 g(7);
    long checked = 0;
    for (uint32_t d = 1; d <= 4096; ++d) {
        qob::FastDiv f = qob::make_fastdiv(d);
        auto chk = [&](uint32_t n) {
            if (n >= (1u << 31)) return true;
            ++checked;
            return qob::fdiv(n, f.m, f.l) == n / d;
        };
        for (uint32_t k = 0; k < 2000; ++k) {
            uint64_t b = (uint64_t)k * d;
            if (!chk((uint32_t)b) || (b && !chk((uint32_t)(b - 1))) || !chk((uint32_t)(b + 1))) {
                printf("FAIL d=%u near %llu\n", d, (unsigned long long)b);
                return 1;
            }
        }
        for (int r = 0; r < 2000; ++r) {
            uint32_t n = (uint32_t)(g() & 0x7fffffffu);
            if (!chk(n)) { printf("FAIL d=%u n=%u\n", d, n); return 1; }
        }
        uint32_t top = 0x7fffffffu;
        if (!chk(top) || !chk(top - d)) { printf("FAIL d=%u top\n", d); return 1; }
    }
    // large strides (products of site dims)
    for (uint32_t d : {6400u, 8192u, 65536u, 1048576u, 1600u * 40u, 3u * 5u * 7u * 11u * 13u * 17u}) {
        qob::FastDiv f = qob::make_fastdiv(d);
        for (int r = 0; r < 200000; ++r) {
            uint32_t n = (uint32_t)(g() & 0x7fffffffu);
            if (qob::fdiv(n, f.m, f.l) != n / d) { printf("FAIL d=%u n=%u\n", d, n); return 1; }
        }
    }
    printf("OK %ld\n", checked);
    return 0;
}

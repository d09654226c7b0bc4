// Host-side QMC / RNG tables (see qmc.h) and the Perlin permutation.
#include "qmc.h"

#include <cstring>

void halton_incremental(int base, uint64_t n, std::vector<float>& out)
{
    out.resize(n);
    const double inv = 1.0 / (double)base;
    double value = 0.0;
    for (uint64_t k = 0; k < n; k++) {
        // Halton::getNext, mcqmc.h:35-49
        const double r = 1.0 - value - 0.0000000001;
        if (inv < r) value += inv;
        else {
            double hh, h = inv;
            do {
                hh = h;
                h *= inv;
            } while (h >= r);
            value += hh + h - 1.0;
        }
        out[k] = (float)value;
    }
}

int next_prime(int last)
{
    int np = last + (last & 1) + 1;
    for (;;) {
        int dv = 3;
        bool ispr = true;
        while (ispr && dv * dv <= np) {
            ispr = (np % dv) != 0;
            dv += 2;
        }
        if (ispr) break;
        np += 2;
    }
    return np;
}

void glibc_rand(uint32_t seed, int64_t n, std::vector<int32_t>& out)
{
    // r[0] = seed; r[i] = 16807 r[i-1] mod (2^31-1) (i < 31); r[i] = r[i-31]
    // for i = 31..33; then r[i] = r[i-31] + r[i-3] mod 2^32; rand() yields
    // r[i] >> 1 from i = 344 on.
    std::vector<uint32_t> r((size_t)(344 + n));
    if (seed == 0) seed = 1;
    r[0] = seed;
    for (int i = 1; i < 31; i++) {
        int64_t v = (16807LL * (int32_t)r[i - 1]) % 2147483647LL;
        if (v < 0) v += 2147483647LL;
        r[i] = (uint32_t)v;
    }
    for (int i = 31; i < 34; i++) r[i] = r[i - 31];
    for (int64_t i = 34; i < 344 + n; i++) r[i] = r[i - 31] + r[i - 3];
    out.resize((size_t)n);
    for (int64_t k = 0; k < n; k++) out[k] = (int32_t)(r[344 + k] >> 1);
}

void lcg_pow_table(uint64_t out[64])
{
    const uint64_t m = 2147483647u;
    uint64_t b = 16807;
    for (int k = 0; k < 64; k++) {
        out[k] = b;
        b = (b * b) % m;
    }
}

// The improved-Perlin permutation of noise.cc:118-127, as hex.
static const char* PERM_HEX =
    "a2a0193bf8ebaaeef31c67281ded00de952edc3f3a82354d6cba36d0f60c7932"
    "d159f4088b63892fb8b49783f28f18c751146587482042a880b54013b2227e57"
    "bc7f6b9d864cc8db7cd5254e5a557450cdb37abbc3cbb6e2e4ecfd980b96d39e"
    "5ca164f18161e1c42472498c904b843438ab78ca1f01d79311c158a931f9446d"
    "bf339c5f0994a38506c69a1e7b461530272b1b713c5bd66f62ac4fc2c00eb123"
    "a7df47b0776905e9e6e776730ffe6e9b56ef12a537fcaed9038edd10b9cec98d"
    "da2abd68179fbed40accd2e8433d70b7027d99d80d608a042c3e92e5af5307e0"
    "29a6c5e3f5f74a41266a165e522d21adf091ffea54fa661a4539cf75a488fb5d";

static int hexv(char c) { return (c <= '9') ? c - '0' : c - 'a' + 10; }

void perm_fill(unsigned char out[512])
{
    for (int i = 0; i < 256; i++) {
        unsigned char v = (unsigned char)(hexv(PERM_HEX[2 * i]) * 16 + hexv(PERM_HEX[2 * i + 1]));
        out[i] = v;
        out[i + 256] = v;
    }
}

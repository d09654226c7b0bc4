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

// The 256 unit gradient vectors of stdPerlin_t (stdp_g, noise.cc:200-370, whose
// 514 rows are these 256 twice plus two) and blenderNoise_t (hashvectf,
// noise.cc:129-150, the same 256 rows), as the bit patterns of their floats.
static const char* GRAD_HEX =
    "3eacf80e3f3737fcbf1c77ffbf71ac04bea737fcbd3ae03bbdceffe3bed53804"
    "bf674bf93f4c9e023efcfbfcbeaf13f0bf5aca043f049e023d0b007a3ed9abf3"
    "bee02ff9bf4ac7fcbeb7c8083f18d8023f37a605bf7c44073e142fe83db6dfc6"
    "bf1a0dfebeacaff7bf3951f8bee65c0c3f1832013f2a98033e555001bddcffeb"
    "3f78d9f93f1347fa3d7780343f50cdfb3e9633f0bf602e02bec453f43e9667f9"
    "3eee77f73f55adfbbf5893fbbe6effe3bef56c0dbded2fe43e127807bf7b9ffd"
    "3e5197e5bf72f406be7547e03f725bfbbe8724103e3cbff4be7140143f12e600"
    "3f48c9fb3e50980bbebe97f2bf67cdff3f41ca08bf26a9f83d6280283efee009"
    "3f508ffcbe983ff7be2608073d830fd0bf7c13fdbea893fb3f5599fe3ee243fe"
    "3f342e02bf06ec07bef3d4063f74d8023e883ff73df54fbe3ef5d3fe3ef737fc"
    "3f3b7803be3ec80c3f7611fdbe500000bf3f9bfebf282bfabdb91ff7bdae2fbe"
    "be2938153f7b8801bf44fbfc3f20d003bdebffce3ec42bfa3f09a4053f4045fe"
    "3f1dd207bf2b2208bed4f3febe851bf8bf2148023f3b53f83f1229fabdb4202e"
    "3f50f3febeea80063cbe7f2c3f637c03be48e0013f7a0008bdb4cfd1be871bef"
    "bf32b7febf2a5ffa3edfd007bf646c01bde70004bf1f2607be6bf7f03f3fa205"
    "3e6e00083f668c05bebc08103e639ffdbf4bd0073f100ffebf2a5e03bf3cd5fa"
    "3de9003f3f2b8df83f346df83e6e30013f6545fa3edc0810bdead03ebdecaff7"
    "bef2cffa3f5f6c053f1f2ffd3f1aa002beff340dbf550bfa3c50fe8b3f0de3fc"
    "bf33d5fe3f34bc06bdb7a030bf3128073f2638003ea17bfabe343809bf7a5e03"
    "3de51fc53dc68fd2bd4adff83f7e78073f22a1f9bf1b4ff8bef49ff9be85c400"
    "bf1b8002bf400a073f7bb7fa3e29680e3d9b8fdebe98c3f43ecee3f8bf5d5a08"
    "bf2380023f01fa013f1402083ec6c3fc3ed2fc053f5301fc3f0e6202be55d80e"
    "3f4dea033f4747fa3f0fe0053e8f1befbf1dd7fe3f482bfabdbb6fe33e48fffc"
    "3d5120193f7aac00bdf840183e541febbf788206be319008bec4fbfc3f6813fd"
    "3e255ffebf3925ffbf2be0053e36c7ebbbcc0379bf7be005be01101b3e21c821"
    "3f7ab6073f549a02bccb40703f0e7408bf029803bf44fa053ec4bc063f507208"
    "3e4ce810bf0b7dfabde68017bec9600fbf6995ff3f0fd8063e1be015bf5023fb"
    "3e4bb01dbe82500dbf724606bed307f23f30e400bf1803ffbed083fd3ea62802"
    "3f5a89fcbf30b0083ebb73f3bf1fc600bedb50093f581a05bea503fbbe595fee"
    "be9857f3bf6e4606be8bcff63f041ffb3f4fd1fe3f6da5fc3ebe47fe3bed02cd"
    "3f492c06bef55c10bec8140ebf3474003d715f893f34f2063e3167ec3e507810"
    "3f76ac003ed82c0abf678002bd7fdf7fbeba0408be3ad81b3f69e2043e820ff5"
    "bf0d63febf4b3ffbbe93ec03bf4407ffbf1315f93d6edfe73e9530063f746df8"
    "3f7235f8be9b9c0f3de3dff4bda010413ee32bfebf648a053e3a77e63f64d5fa"
    "3ed1abf3bf2e2802be5a70093f337dfa3d7fdf7f3f593006bf069004bf3a91fb"
    "bf2450053e740fedbf6182023edaf7f53e4f8013bf0851fcbe5b2fec3f51a1fd"
    "3f72f007be9c33f83da30fd03f6cea073ebfd3f63d64a05e3f5228023f10c405"
    "3da2003b3e84840e3dcc5fc0bf75f202be96d80a3f6917f83e9477ff3e0c67e0"
    "3ea3dc05bf6ff9f8bf688bfc3eb0c804be71600fbf6fabfc3e0e17e33ea54003"
    "3ebbaffb3f6d0a03bdb99fe43e9e5c0cbf2fd406bf285dfb3f2b79fb3cc840e1"
    "3f3df7f9bf6b01fcbd722079bec8bbf53ebb23ff3eec9ff9bf4ed3fa3daabffd"
    "3f7f1401bc727fe53f1806083e819bf6bf4380023f799807be296feb3e17d806"
    "3f6b42033e0c981c3ebd33f43e9703f33f31ea033f27ddfd3f716e053e1c47e5"
    "be97340d3f167c07bf1925ff3f0b64063e50afe63f2da0063f34bc06bef52807"
    "bf2938043f13e1fc3f3813fdbe2de00dbf2c7ffebe075819bf771fff3e6667f9"
    "bf7b5c07be14801f3df9d0203f4371fc3f1469f93e91b003bf64a7f83da32fcb"
    "3ee28bf4bf4bb7fabf0618053e9b8bf03f3c26033f2d9408bbf3fd933e1ec80c"
    "3f7c7bfebd6620253e841bfc3f124e05bf476806bf2e79ffbf33bbf9be531013"
    "bf767dfebe356ffc3e507fedbef12407bf03ddfd3f375200bf3d86013f1243fe"
    "3eb54bf1bf0f6406bf533205bd99902abf0204083f40d1fa3ed60008bf00cc03"
    "3d1c40073f5d03fb3f1812063f2c6df8bee127f6bdf57ffabbb502acbf7e2403"
    "3dc94035be5ae8193f78cdfbbf1d99fe3f24d9f93ee8880e3f659807bee1d40f"
    "3d03c07fbf0e15fd3f403201beb757f73ecc2bfa3e9c00113f5d6bfdbf3913f9"
    "3e9b70043f1edbf9bf22b607bf1f37fcbeea2403be9623f23d94202e3f740a07"
    "befbb7fa3f2939fcbf1117f8be9bbff9bf127408bf42fdfc3f6877ffbecc43f6"
    "3e020ff5bf060c07bf0cb806bf269df9bf6ed3fabe4ca81a3e994c013dcb9fdc"
    "3e103ff73f7c27fabca6809dbf7ff002bb17fa5d3f73e1fc3e8fcbf7bdede00d"
    "bf78cdfb3da9203a3e61a0063f28c1fd3f348e04be85740c3f3baa04be33bff9"
    "3f2833f8bf0e21fb3edbe7f3bf3652053ecd03fbbf6609ff3e37a80d3f403a00"
    "be49afe23f215e073f4921ffbf11c0013e779fedbd714014beed9ff53f623fff"
    "3e04b7f6bf71f7f93e99680ebeb7340d3f685c033e607fedbf579607bee0d3f2"
    "bea037f8bea0800f3f5826033ede7408be7700043d5a1fd13f780ffe3d82cfd9"
    "3f52c2053f1060033ef3dc053f5cb5febe30a8093f1ef5fdbe98a3f93f399408"
    "3e8137f4bf3fd5febf1cb407bf13d9fd3ec5ebfb3f380ffebed00c0fbf3cebff"
    "3f09e7fb3d413fd13f0e6a013f546003bf4d87f93f1663fa3dcf301fbf352dfd"
    "bf34b9fe3cd87f883ebff3f13f2f3ffb3f2006083ef1bbf13e55681fbf5b43fa"
    "bf33fc01bf14c1fdbed1d40fbed59c06bf3c7ffe3f0859fbbee55001be4fb7e9"
    "bf5ee8083db21038bf7d8bf83ddb4fe83e3c17ec3c963fddbf7b97ff3f0f85f9"
    "bedb94103f3551f83e97ebf2bf73dc05bd89a0273e0b08143f7d82023cf7c02b"
    "bd8ce0333f279bfe3f40b0083f433a04bf0d40033eacd7f1bf51d207bed0bbf5"
    "3ece1006bf3a53fcbf0cfbfcbed12c06bf5b1401bef5c4003e456ffc3f3173fb"
    "bda2602d3f3761fe3e676ff43f266e01bf39b5fa3e7c9ff93f740ffebe31b7e1"
    "bf786c01be74f7ecbd04dfce3e0ba7fcbec9e40d3f68a3f93f502ff93f0ee408"
    "3e2850163ecf1bef3f0c8e04bf3b37fcbec2afffbf1113f93f3b19f83cb58080"
    "3f6ec7fc3eb82ff93c0701113f7f1e08bda8900c3f74e608bd85cfeebe914403"
    "bf3e5a043c067f0b3f2b27febf5c0e06bf0279ff3d12806bbf5a2c023eba3401"
    "bec0800fbe4ecfeabf788605be043fe63ea0800f3f7284063d860fe4be820ff5"
    "bf073bfb3f4f65fe3f20c8043f19e2043efcfbfcbefd20083f5e4e053d1b5fa2"
    "3e8d2008bdb1001d3f7513f9be4aafff3ef0140ebf5c59fb3f6545fa3ecc7c0f"
    "3e48800f3f0f90003f5333fdbd8dc0123e9bbff9bf5977fb3edcaff73dc56042"
    "3f6f8df83ead9ff5bd529fe03ed163febf693c043ea92ff53f715e07bd2da017"
    "bee79bfebf4951f83ed754093e098809bf6ef805beaa3c003f4d49f9bf1113f9"
    "be4177eabf7cea07bdd9103d3de66fd6be79f01c3f6469f9bec273f7bf6ba605"
    "3e6a97e1bea1fc093f4765fe3ea6e3f83f092407bf69b1fb3e3da7f4beba4010"
    "be3d38043f10bc06bf4dc800bc973fb83df4cfd13f7e1c04be84bc06bf6f8402"
    "3e7547e0bf52dff8bee5f3f9beb11ff7beddbff03f28d5fa3f1d45fabf52a002"
    "3ec1f7f9bed8ec103f2fec03bf2cc1fdbe89abf3be7c7ffebf4d03fb3f0bb3fa"
    "bf3aac00be54c7f33f26e8083f274e01bf1c45febee5100a3efc33f8bcbf401c"
    "3f5eb3ff3f1be9fb3c14ff00bf4b05fb3f766602be8acc0cbc10ff973e88240b"
    "bba0fd3b3f76c6043f26b7fe3f0db9fabf04d9f93e8fa3fdbf57d8063eeacffa"
    "be338002bf7bb3fa3d4c00543d10bfa1bf7adbf93e48c005bf7ba2053e284018"
    "bda870113e6f57f7bf78e608bba400fcbf3f4bf9bf1c7ffe3e8558003f2545fa"
    "3f17bff93ef677f73df10fd8bf731401be94d402bf08f606bf4aba06be96b010"
    "bf208e04be5b97ff3f3fabfcbd448063bf504007bf146003be33bff93ef45bf2"
    "bf5c69f93f3d1202bed4180dbf0826033f1003ffbf343dff3eddd80ebf414403"
    "3f25e3fcbdce00083deaf0393d365fd9bf7e0c073f774bf93e7abffdbda9802c";

void grad_fill(float out[768])
{
    for (int i = 0; i < 768; i++) {
        uint32_t b = 0;
        for (int k = 0; k < 8; k++) b = b * 16 + (uint32_t)hexv(GRAD_HEX[8 * i + k]);
        memcpy(&out[i], &b, 4);
    }
}

/*
 * CPU oracle: improved Perlin noise, fBm, turbulence and the procedural
 * textures built on them.  TEST INFRASTRUCTURE ONLY (see yaf_oracle.h).
 *
 * Transcendentals whose reference call is a float libm function (sinf, powf)
 * are evaluated as (float)f((double)x) here and on the device, so both sides
 * agree bit for bit; they differ from libm's float functions by at most one
 * unit in the last place.
 */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include "yo_internal.h"

/* The 256-entry permutation of noise.cc:118-127 (the 512-entry table is this
 * permutation twice).  Kept as data: the noise values depend on it exactly. */
static const char* PERM_HEX =
    "a2a0193bf8ebaaeef31c67281ded00de952edc3f3a82354d6cba36d0f60c7932"
    "d159f4088b63892fb8b49783f28f18c751146587482042a880b54013b2227e57"
    "bc7f6b9d864cc8db7cd5254e5a557450cdb37abbc3cbb6e2e4ecfd980b96d39e"
    "5ca164f18161e1c42472498c904b843438ab78ca1f01d79311c158a931f9446d"
    "bf339c5f0994a38506c69a1e7b461530272b1b713c5bd66f62ac4fc2c00eb123"
    "a7df47b0776905e9e6e776730ffe6e9b56ef12a537fcaed9038edd10b9cec98d"
    "da2abd68179fbed40accd2e8433d70b7027d99d80d608a042c3e92e5af5307e0"
    "29a6c5e3f5f74a41266a165e522d21adf091ffea54fa661a4539cf75a488fb5d";

static unsigned char HASH[512];
static pthread_once_t hash_once = PTHREAD_ONCE_INIT;

static int hexv(char c) { return (c <= '9') ? c - '0' : c - 'a' + 10; }

static void hash_fill(void)
{
    for (int i = 0; i < 256; i++) {
        unsigned char v = (unsigned char)(hexv(PERM_HEX[2 * i]) * 16 + hexv(PERM_HEX[2 * i + 1]));
        HASH[i] = v;
        HASH[i + 256] = v;
    }
}

/* renders on several threads share the tables: filled once */
static void hash_init(void) { pthread_once(&hash_once, hash_fill); }

void yo_perm(unsigned char* out)
{
    hash_init();
    memcpy(out, HASH, 512);
}

static inline float dsin(float x) { return (float)sin((double)x); }
static inline float dpow(float x, float y) { return (float)pow((double)x, (double)y); }

/* newPerlin_t::fade / grad, noise.h:33-40 */
static inline float fade(float t) { return t * t * t * (t * (t * 6 - 15) + 10); }
static inline float grad(int hash, float x, float y, float z)
{
    int h = hash & 15;
    float u = h < 8 ? x : y, v = h < 4 ? y : h == 12 || h == 14 ? x : z;
    return ((h & 1) == 0 ? u : -u) + ((h & 2) == 0 ? v : -v);
}
static inline float lerpf(float t, float a, float b) { return a + t * (b - a); }

/* newPerlin_t::operator(), noise.cc:161-183 */
float yo_newperlin(float x, float y, float z)
{
    hash_init();
    float u = floorf(x), v = floorf(y), w = floorf(z);
    int X = ((int)u) & 255, Y = ((int)v) & 255, Z = ((int)w) & 255;
    x -= u; y -= v; z -= w;
    u = fade(x); v = fade(y); w = fade(z);
    int A = HASH[X] + Y, AA = HASH[A] + Z, AB = HASH[A + 1] + Z;
    int B = HASH[X + 1] + Y, BA = HASH[B] + Z, BB = HASH[B + 1] + Z;
    float nv = lerpf(w,
                     lerpf(v, lerpf(u, grad(HASH[AA], x, y, z), grad(HASH[BA], x - 1, y, z)),
                           lerpf(u, grad(HASH[AB], x, y - 1, z), grad(HASH[BB], x - 1, y - 1, z))),
                     lerpf(v, lerpf(u, grad(HASH[AA + 1], x, y, z - 1), grad(HASH[BA + 1], x - 1, y, z - 1)),
                           lerpf(u, grad(HASH[AB + 1], x, y - 1, z - 1),
                                 grad(HASH[BB + 1], x - 1, y - 1, z - 1))));
    return (float)(0.5 + 0.5 * (double)nv);
}

/* The 256 unit gradients of stdPerlin_t (stdp_g, noise.cc:200-370: these
 * rows twice plus two) and blenderNoise_t (hashvectf, noise.cc:129-150: the
 * same rows), as the bit patterns of their floats. */
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
static float GRAD[768];
static pthread_once_t grad_once = PTHREAD_ONCE_INIT;

static void grad_fill(void)
{
    for (int i = 0; i < 768; i++) {
        uint32_t b = 0;
        for (int k = 0; k < 8; k++) b = b * 16 + (uint32_t)hexv(GRAD_HEX[8 * i + k]);
        memcpy(&GRAD[i], &b, 4);
    }
}

static void grad_init(void) { pthread_once(&grad_once, grad_fill); }

void yo_grad_table(float* out)
{
    grad_init();
    memcpy(out, GRAD, sizeof(GRAD));
}

/* stdPerlin_t::operator(), noise.cc:372-435.  setup(): t = v + 10000.0 is a
 * double sum stored in a float; surve(t) = t*t*(3.0-2.0*t) is formed in
 * double; the result 0.5+0.75*lerp is double.  Rows of stdp_p / stdp_g past
 * 255 repeat the first ones. */
float yo_stdperlin(float x, float y, float z)
{
    hash_init();
    grad_init();
    const float vec[3] = {x, y, z};
    int b0[3], b1[3];
    float r0[3], r1[3], s[3];
    for (int k = 0; k < 3; k++) {
        float t = (float)((double)vec[k] + 10000.0);
        b0[k] = ((int)t) & 255;
        b1[k] = (b0[k] + 1) & 255;
        r0[k] = t - (float)(int)t;
        r1[k] = (float)((double)r0[k] - 1.0);
        s[k] = (float)((double)(r0[k] * r0[k]) * (3.0 - 2.0 * (double)r0[k]));
    }
    int i = HASH[b0[0]], j = HASH[b1[0]];
    int b00 = HASH[i + b0[1]], b10 = HASH[j + b0[1]], b01 = HASH[i + b1[1]], b11 = HASH[j + b1[1]];
#define STDP_AT(row, rx, ry, rz) ((rx) * GRAD[3 * ((row) & 255)] + (ry) * GRAD[3 * ((row) & 255) + 1] + \
                                  (rz) * GRAD[3 * ((row) & 255) + 2])
    float u, v, a, b, c, d;
    u = STDP_AT(b00 + b0[2], r0[0], r0[1], r0[2]);
    v = STDP_AT(b10 + b0[2], r1[0], r0[1], r0[2]);
    a = lerpf(s[0], u, v);
    u = STDP_AT(b01 + b0[2], r0[0], r1[1], r0[2]);
    v = STDP_AT(b11 + b0[2], r1[0], r1[1], r0[2]);
    b = lerpf(s[0], u, v);
    c = lerpf(s[1], a, b);
    u = STDP_AT(b00 + b1[2], r0[0], r0[1], r1[2]);
    v = STDP_AT(b10 + b1[2], r1[0], r0[1], r1[2]);
    a = lerpf(s[0], u, v);
    u = STDP_AT(b01 + b1[2], r0[0], r1[1], r1[2]);
    v = STDP_AT(b11 + b1[2], r1[0], r1[1], r1[2]);
    b = lerpf(s[0], u, v);
    d = lerpf(s[1], a, b);
#undef STDP_AT
    return (float)(0.5 + 0.75 * (double)lerpf(s[2], c, d));
}

/* blenderNoise_t::operator(), noise.cc:439-509: weights 1-3t^2+2t^3 (ox)
 * and 1-3t^2-2t^3 (jx = ox-1) in double, n summed in float from 0.5,
 * clamped to [0, 1]. */
float yo_blendernoise(float x, float y, float z)
{
    hash_init();
    grad_init();
    int ix = (int)floor(x), iy = (int)floor(y), iz = (int)floor(z);
    float ox = x - (float)ix, oy = y - (float)iy, oz = z - (float)iz;
    float jx = ox - 1, jy = oy - 1, jz = oz - 1;
    float cn1 = ox * ox, cn2 = oy * oy, cn3 = oz * oz, cn4 = jx * jx, cn5 = jy * jy, cn6 = jz * jz;
    cn1 = (float)(1.0 - 3.0 * cn1 + 2.0 * cn1 * ox);
    cn2 = (float)(1.0 - 3.0 * cn2 + 2.0 * cn2 * oy);
    cn3 = (float)(1.0 - 3.0 * cn3 + 2.0 * cn3 * oz);
    cn4 = (float)(1.0 - 3.0 * cn4 - 2.0 * cn4 * jx);
    cn5 = (float)(1.0 - 3.0 * cn5 - 2.0 * cn5 * jy);
    cn6 = (float)(1.0 - 3.0 * cn6 - 2.0 * cn6 * jz);
    int b00 = HASH[HASH[ix & 255] + (iy & 255)], b10 = HASH[HASH[(ix + 1) & 255] + (iy & 255)];
    int b01 = HASH[HASH[ix & 255] + ((iy + 1) & 255)], b11 = HASH[HASH[(ix + 1) & 255] + ((iy + 1) & 255)];
    int b20 = iz & 255, b21 = (iz + 1) & 255;
    const float* h;
    float n = 0.5f, i;
    i = cn1 * cn2 * cn3; h = GRAD + 3 * HASH[b20 + b00]; n += i * (h[0] * ox + h[1] * oy + h[2] * oz);
    i = cn1 * cn2 * cn6; h = GRAD + 3 * HASH[b21 + b00]; n += i * (h[0] * ox + h[1] * oy + h[2] * jz);
    i = cn1 * cn5 * cn3; h = GRAD + 3 * HASH[b20 + b01]; n += i * (h[0] * ox + h[1] * jy + h[2] * oz);
    i = cn1 * cn5 * cn6; h = GRAD + 3 * HASH[b21 + b01]; n += i * (h[0] * ox + h[1] * jy + h[2] * jz);
    i = cn4 * cn2 * cn3; h = GRAD + 3 * HASH[b20 + b10]; n += i * (h[0] * jx + h[1] * oy + h[2] * oz);
    i = cn4 * cn2 * cn6; h = GRAD + 3 * HASH[b21 + b10]; n += i * (h[0] * jx + h[1] * oy + h[2] * jz);
    i = cn4 * cn5 * cn3; h = GRAD + 3 * HASH[b20 + b11]; n += i * (h[0] * jx + h[1] * jy + h[2] * oz);
    i = cn4 * cn5 * cn6; h = GRAD + 3 * HASH[b21 + b11]; n += i * (h[0] * jx + h[1] * jy + h[2] * jz);
    if (n < 0.0f) n = 0.0f;
    else if (n > 1.0f) n = 1.0f;
    return n;
}

/* the texture's noise generator (newNoise, basictex.cc:38-63) */
static float noise_gen(int kind, float x, float y, float z)
{
    if (kind == YG_NOISE_STDPERLIN) return yo_stdperlin(x, y, z);
    if (kind == YG_NOISE_BLENDER) return yo_blendernoise(x, y, z);
    return yo_newperlin(x, y, z);
}

/* getSignedNoise, noise.h:245-248 */
static inline float signed_noise(int kind, float x, float y, float z)
{
    return 2.0f * noise_gen(kind, x, y, z) - 1.0f;
}

/* fBm_t::operator(), noise.cc:638-650 */
static float fbm_k(int kind, float x, float y, float z, float H, float lac, float oct)
{
    float value = 0, pwr = 1, pwHL = dpow(lac, -H);
    for (int i = 0; i < (int)oct; i++) {
        value += signed_noise(kind, x, y, z) * pwr;
        pwr *= pwHL;
        x *= lac; y *= lac; z *= lac;
    }
    float rmd = oct - floorf(oct);
    if (rmd != 0.f) value += rmd * signed_noise(kind, x, y, z) * pwr;
    return value;
}

/* mFractal_t::operator(), noise.cc:668-681 */
static float mfractal_k(int kind, float x, float y, float z, float H, float lac, float oct)
{
    float value = 1, pwr = 1, pwHL = dpow(lac, -H);
    for (int i = 0; i < (int)oct; i++) {
        value *= (pwr * signed_noise(kind, x, y, z) + 1.0f);
        pwr *= pwHL;
        x *= lac; y *= lac; z *= lac;
    }
    float rmd = oct - floorf(oct);
    if (rmd != 0.0f) value *= (rmd * signed_noise(kind, x, y, z) * pwr + 1.0f);
    return value;
}

/* heteroTerrain_t::operator(), noise.cc:695-719 */
static float hetero_k(int kind, float x, float y, float z, float H, float lac, float oct, float offset)
{
    float pwHL = dpow(lac, -H), pwr = pwHL;
    float value = offset + signed_noise(kind, x, y, z), increment;
    x *= lac; y *= lac; z *= lac;
    for (int i = 1; i < (int)oct; i++) {
        increment = (signed_noise(kind, x, y, z) + offset) * pwr * value;
        value += increment;
        pwr *= pwHL;
        x *= lac; y *= lac; z *= lac;
    }
    float rmd = oct - floorf(oct);
    if (rmd != 0.0f) {
        increment = (signed_noise(kind, x, y, z) + offset) * pwr * value;
        value += rmd * increment;
    }
    return value;
}

/* hybridMFractal_t::operator(), noise.cc:731-754 */
static float hybrid_k(int kind, float x, float y, float z, float H, float lac, float oct, float offset, float gain)
{
    float pwHL = dpow(lac, -H), pwr = pwHL;
    float result = signed_noise(kind, x, y, z) + offset;
    float weight = gain * result;
    x *= lac; y *= lac; z *= lac;
    for (int i = 1; (weight > 0.001f) && (i < (int)oct); i++) {
        if (weight > 1.0f) weight = 1.0f;
        float signal = (signed_noise(kind, x, y, z) + offset) * pwr;
        pwr *= pwHL;
        result += weight * signal;
        weight *= gain * signal;
        x *= lac; y *= lac; z *= lac;
    }
    float rmd = oct - floorf(oct);
    if (rmd != 0.0f) result += rmd * ((signed_noise(kind, x, y, z) + offset) * pwr);
    return result;
}

/* ridgedMFractal_t::operator(), noise.cc:768-791 */
static float ridged_k(int kind, float x, float y, float z, float H, float lac, float oct, float offset, float gain)
{
    float pwHL = dpow(lac, -H), pwr = pwHL;
    float signal = offset - fabsf(signed_noise(kind, x, y, z));
    signal *= signal;
    float result = signal, weight = 1.0f;
    for (int i = 1; i < (int)oct; i++) {
        x *= lac; y *= lac; z *= lac;
        weight = signal * gain;
        if (weight > 1.0f) weight = 1.0f;
        else if (weight < 0.0f) weight = 0.0f;
        signal = offset - fabsf(signed_noise(kind, x, y, z));
        signal *= signal;
        signal *= weight;
        result += signal * pwr;
        pwr *= pwHL;
    }
    return result;
}

/* the generator textureMusgrave_t's "type" selects (basictex.cc:657-672) */
static float musgrave_k(const yg_texture* t, float x, float y, float z)
{
    switch (t->mtype) {
    case YG_MUSGRAVE_MULTIFRACTAL: return mfractal_k(t->noise, x, y, z, t->H, t->lacunarity, t->octaves);
    case YG_MUSGRAVE_HETEROTERRAIN: return hetero_k(t->noise, x, y, z, t->H, t->lacunarity, t->octaves, t->offset);
    case YG_MUSGRAVE_HYBRIDMF:
        return hybrid_k(t->noise, x, y, z, t->H, t->lacunarity, t->octaves, t->offset, t->gain);
    case YG_MUSGRAVE_RIDGEDMF:
        return ridged_k(t->noise, x, y, z, t->H, t->lacunarity, t->octaves, t->offset, t->gain);
    default: return fbm_k(t->noise, x, y, z, t->H, t->lacunarity, t->octaves);
    }
}

float yo_fbm(float x, float y, float z, float H, float lac, float oct)
{
    return fbm_k(YG_NOISE_NEWPERLIN, x, y, z, H, lac, oct);
}

/* turbulence, noise.cc:808-819: only blenderNoise_t offsets the point by
 * (1, 1, 1) (noise.h:60-61) before the size scale */
static float turbulence_k(int kind, float x, float y, float z, int oct, float size, int hard)
{
    float val, amp = 1, sum = 0;
    if (kind == YG_NOISE_BLENDER) { x += 1.0f; y += 1.0f; z += 1.0f; }
    x *= size; y *= size; z *= size;
    for (int i = 0; i <= oct; i++, amp *= 0.5f, x *= 2.0f, y *= 2.0f, z *= 2.0f) {
        val = noise_gen(kind, x, y, z);
        if (hard) val = (float)fabs(2.0 * (double)val - 1.0);
        sum += amp * val;
    }
    return sum * ((float)(1 << oct) / (float)((1 << (oct + 1)) - 1));
}

float yo_turbulence(float x, float y, float z, int oct, float size, int hard)
{
    return turbulence_k(YG_NOISE_NEWPERLIN, x, y, z, oct, size, hard);
}

/* texture getFloat: clouds basictex.cc:82-90, marble :133-153, wood :194-216,
 * Musgrave :677-680 */
static float tex_float(const yg_texture* t, yo_v3 p)
{
    switch (t->type) {
    case YG_TEX_CLOUDS: {
        float v = turbulence_k(t->noise, p.x, p.y, p.z, t->depth, t->size, t->hard);
        if (t->bias) {
            v *= v;
            if (t->bias == 1) return -v;
        }
        return v;
    }
    case YG_TEX_MARBLE:
    case YG_TEX_WOOD: {
        float w;
        if (t->type == YG_TEX_MARBLE) {
            w = (float)((double)(p.x + p.y + p.z) * 5.0 +
                        ((t->turb == 0.0f) ? 0.0 : (double)(t->turb * turbulence_k(t->noise, p.x, p.y, p.z, t->depth, t->size, t->hard))));
        } else {
            if (t->rings) w = (float)((double)sqrtf(p.x * p.x + p.y * p.y + p.z * p.z) * 20.0);
            else w = (float)((double)(p.x + p.y + p.z) * 10.0);
            w = (float)((double)w + ((t->turb == 0.0f) ? 0.0 : (double)(t->turb * turbulence_k(t->noise, p.x, p.y, p.z, t->depth, t->size, t->hard))));
        }
        switch (t->shape) {
        case YG_SHAPE_SAW:
            w *= (float)(0.5 * M_1_PI);
            w -= floorf(w);
            break;
        case YG_SHAPE_TRI:
            w *= (float)(0.5 * M_1_PI);
            w = fabsf(2.0f * (w - floorf(w)) - 1.0f);
            break;
        default:
            w = 0.5f + 0.5f * dsin(w);
        }
        if (t->type == YG_TEX_MARBLE) return dpow(w, t->sharpness);
        return w;
    }
    case YG_TEX_MUSGRAVE:
        /* textureMusgrave_t::getFloat (basictex.cc:677-680) */
        return t->iscale * musgrave_k(t, p.x * t->size, p.y * t->size, p.z * t->size);
    }
    return 0.0f;
}

/* getColor: color1 + getFloat(p)*(color2 - color1) (basictex.cc:92-95 etc.) */
static yo_col tex_color(const yg_texture* t, yo_v3 p)
{
    float f = tex_float(t, p);
    yo_col c1 = col(t->color1[0], t->color1[1], t->color1[2]);
    yo_col c2 = col(t->color2[0], t->color2[1], t->color2[2]);
    return col_add(c1, col_scale(col_sub(c2, c1), f));
}

/* modulator_t::doMapping, procedural branch (texture.cc:167-180, 311-317)
 * from the texco's coordinates P: the texture matrix when given
 * (matrix4x4_t * point3d_t, matrix4.h:86-91; for a direction, vec != 0, the
 * vector form without translation, matrix4.h:79-84), then
 * size * (T[projected axis] + ofs) per axis, or size * ofs for axis "n". */
void yo_mod_texpt_v(const yg_modulator* m, yo_v3 P, int vec, float out[3])
{
    float T[3] = {P.x, P.y, P.z};
    if (m->has_mtx) {
        const float* M = m->mtx;
        const float tx = vec ? 0.0f : M[3], ty = vec ? 0.0f : M[7], tz = vec ? 0.0f : M[11];
        T[0] = M[0] * P.x + M[1] * P.y + M[2] * P.z;
        T[1] = M[4] * P.x + M[5] * P.y + M[6] * P.z;
        T[2] = M[8] * P.x + M[9] * P.y + M[10] * P.z;
        if (!vec) {
            T[0] = T[0] + tx;
            T[1] = T[1] + ty;
            T[2] = T[2] + tz;
        }
    }
    for (int a = 0; a < 3; a++) {
        const int pj = m->proj[a];
        out[a] = (pj < 0) ? m->size[a] * m->ofs[a] : m->size[a] * (T[pj == 0 ? a : pj - 1] + m->ofs[a]);
    }
}

void yo_mod_texpt(const yg_modulator* m, yo_v3 P, float out[3]) { yo_mod_texpt_v(m, P, 0, out); }

/* modulator_t::modulate colour part (texture.cc:499-534) at the texco's
 * coordinates P (vec: a direction) */
yo_col yo_modulate_v(const yg_modulator* m, const yg_texture* texs, yo_col C, yo_v3 P, int vec)
{
    const yg_texture* t = &texs[m->tex];
    float q[3];
    yo_mod_texpt_v(m, P, vec, q);
    yo_v3 tp = v3(q[0], q[1], q[2]);
    yo_col tc = tex_color(t, tp);
    switch (m->mode) {
    case YG_MOD_MIX:
        if (m->color > 0) C = col_mix(tc, C, m->color);
        break;
    case YG_MOD_MUL:
        if (m->color > 0) C = col_mul(col_mix(tc, col(1, 1, 1), m->color), C);
        break;
    case YG_MOD_ADD:
        if (m->color > 0) C = col_add(col_scale(tc, m->color), C);
        break;
    case YG_MOD_SUB:
        if (m->color > 0) C = col_add(col_scale(tc, -m->color), C);
        break;
    }
    return C;
}

yo_col yo_modulate(const yg_modulator* m, const yg_texture* texs, yo_col C, yo_v3 P)
{
    return yo_modulate_v(m, texs, C, P, 0);
}

void yo_diffuse_at(const yg_material* mats, const yg_modulator* mods, const yg_texture* texs, int mat,
                   const float* pts, int64_t n, float* out)
{
    const yg_material* m = &mats[mat];
    for (int64_t i = 0; i < n; i++) {
        yo_col c = col(m->color[0], m->color[1], m->color[2]);
        for (int k = 0; k < m->mod_count; k++) c = yo_modulate(&mods[m->mod_first + k], texs, c, v3_ld(pts + 3 * i));
        out[3 * i] = c.r; out[3 * i + 1] = c.g; out[3 * i + 2] = c.b;
    }
}

int yo_noise_eval(int what, const yg_texture* tex, const float* pts, int64_t n, float* out)
{
    for (int64_t i = 0; i < n; i++) {
        yo_v3 p = v3_ld(pts + 3 * i);
        switch (what) {
        case YG_EVAL_NOISE: out[i] = noise_gen(tex ? tex->noise : YG_NOISE_NEWPERLIN, p.x, p.y, p.z); break;
        case YG_EVAL_FBM: out[i] = fbm_k(tex->noise, p.x, p.y, p.z, tex->H, tex->lacunarity, tex->octaves); break;
        case YG_EVAL_TURBULENCE:
            out[i] = turbulence_k(tex->noise, p.x, p.y, p.z, tex->depth, tex->size, tex->hard);
            break;
        case YG_EVAL_TEX_FLOAT: out[i] = tex_float(tex, p); break;
        case YG_EVAL_TEX_COLOR: {
            yo_col c = tex_color(tex, p);
            out[3 * i] = c.r; out[3 * i + 1] = c.g; out[3 * i + 2] = c.b;
            break;
        }
        default: return -1;
        }
    }
    return 0;
}

/* Heightfield: x = ex*i/(nx-1), y = ey*j/(ny-1), z = amp*fBm(freq*(x,y,0)) */
void yo_gen_heightfield(int nx, int ny, float ex, float ey, float amp, float freq, float H,
                        float lac, float oct, float* xyz)
{
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            float x = ex * (float)i / (float)(nx - 1);
            float y = ey * (float)j / (float)(ny - 1);
            float z = amp * yo_fbm(freq * x, freq * y, 0.0f, H, lac, oct);
            int64_t k = (int64_t)j * nx + i;
            xyz[3 * k] = x; xyz[3 * k + 1] = y; xyz[3 * k + 2] = z;
        }
}

/* ---------------------------------------------- FP64 texture derivatives -- */

static double fade_d(double t) { return 30.0 * t * t * (t - 1.0) * (t - 1.0); }
static double fade_dd(double t) { return t * t * t * (t * (t * 6 - 15) + 10); }

static void grad_vec(int hash, double g[3])
{
    int h = hash & 15;
    double su = (h & 1) ? -1.0 : 1.0, sv = (h & 2) ? -1.0 : 1.0;
    g[0] = g[1] = g[2] = 0;
    if (h < 8) g[0] += su; else g[1] += su;
    if (h < 4) g[1] += sv;
    else if (h == 12 || h == 14) g[0] += sv;
    else g[2] += sv;
}

/* d newPerlin_t / d(x,y,z) in double: the noise is a trilinear blend (fade
 * weights) of the 8 corner gradient ramps, differentiated analytically */
static void perlin_grad_d(double x, double y, double z, double out[3])
{
    hash_init();
    double fx = floor(x), fy = floor(y), fz = floor(z);
    int X = ((int)fx) & 255, Y = ((int)fy) & 255, Z = ((int)fz) & 255;
    x -= fx; y -= fy; z -= fz;
    double u = fade_dd(x), v = fade_dd(y), w = fade_dd(z);
    double du = fade_d(x), dv = fade_d(y), dw = fade_d(z);
    int A = HASH[X] + Y, AA = HASH[A] + Z, AB = HASH[A + 1] + Z;
    int B = HASH[X + 1] + Y, BA = HASH[B] + Z, BB = HASH[B + 1] + Z;
    int hc[8] = {HASH[AA], HASH[BA], HASH[AB], HASH[BB], HASH[AA + 1], HASH[BA + 1], HASH[AB + 1], HASH[BB + 1]};
    double g[8], G[8][3];
    for (int k = 0; k < 8; k++) {
        double c[3] = {x - (k & 1), y - ((k >> 1) & 1), z - ((k >> 2) & 1)};
        grad_vec(hc[k], G[k]);
        g[k] = G[k][0] * c[0] + G[k][1] * c[1] + G[k][2] * c[2];
    }
    double k1 = g[1] - g[0], k2 = g[2] - g[0], k3 = g[4] - g[0];
    double k4 = g[0] - g[1] - g[2] + g[3], k5 = g[0] - g[2] - g[4] + g[6], k6 = g[0] - g[1] - g[4] + g[5];
    double k7 = -g[0] + g[1] + g[2] - g[3] + g[4] - g[5] - g[6] + g[7];
    double d[3] = {du * (k1 + k4 * v + k6 * w + k7 * v * w), dv * (k2 + k4 * u + k5 * w + k7 * u * w),
                   dw * (k3 + k5 * v + k6 * u + k7 * u * v)};
    for (int k = 0; k < 8; k++) {
        double wu = (k & 1) ? u : 1 - u, wv = ((k >> 1) & 1) ? v : 1 - v, ww = ((k >> 2) & 1) ? w : 1 - w;
        for (int a = 0; a < 3; a++) d[a] += G[k][a] * wu * wv * ww;
    }
    for (int a = 0; a < 3; a++) out[a] = 0.5 * d[a];
}

/* d getFloat(p) / d lambda for p scaled by lambda at 1 (= grad . p), for
 * Musgrave fBm and clouds turbulence */
double yo_tex_dF_dlambda(const yg_texture* t, const float* pp)
{
    double p[3] = {pp[0], pp[1], pp[2]}, g[3] = {0, 0, 0}, q[3];
    if (t->type == YG_TEX_MUSGRAVE) {
        double s = t->size, pwr = 1, scale = 1, pwHL = pow((double)t->lacunarity, -(double)t->H);
        for (int i = 0; i < (int)t->octaves; i++) {
            perlin_grad_d(p[0] * s * scale, p[1] * s * scale, p[2] * s * scale, q);
            for (int a = 0; a < 3; a++) g[a] += 2.0 * pwr * scale * q[a];
            pwr *= pwHL;
            scale *= t->lacunarity;
        }
        double rmd = t->octaves - floor(t->octaves);
        if (rmd != 0) {
            perlin_grad_d(p[0] * s * scale, p[1] * s * scale, p[2] * s * scale, q);
            for (int a = 0; a < 3; a++) g[a] += 2.0 * rmd * pwr * scale * q[a];
        }
        for (int a = 0; a < 3; a++) g[a] *= t->iscale * s;
    } else if (t->type == YG_TEX_CLOUDS) {
        double s = t->size, amp = 1, scale = 1, v = 0;
        for (int i = 0; i <= t->depth; i++, amp *= 0.5, scale *= 2) {
            double x = p[0] * s * scale, y = p[1] * s * scale, z = p[2] * s * scale;
            double n = yo_newperlin((float)x, (float)y, (float)z);
            perlin_grad_d(x, y, z, q);
            double f = scale;
            double val = n;
            if (t->hard) {
                val = fabs(2 * n - 1);
                f *= (2 * n - 1) >= 0 ? 2.0 : -2.0;
            }
            v += amp * val;
            for (int a = 0; a < 3; a++) g[a] += amp * f * q[a];
        }
        double norm = (double)(1 << t->depth) / (double)((1 << (t->depth + 1)) - 1);
        v *= norm;
        for (int a = 0; a < 3; a++) g[a] *= norm * s;
        if (t->bias) for (int a = 0; a < 3; a++) g[a] *= (t->bias == 1) ? -2.0 * v : 2.0 * v;
    }
    return g[0] * p[0] + g[1] * p[1] + g[2] * p[2];
}

void yo_tex_color_f(const yg_texture* t, const float* p, float* out)
{
    yo_col c = tex_color(t, v3_ld(p));
    out[0] = c.r; out[1] = c.g; out[2] = c.b;
}

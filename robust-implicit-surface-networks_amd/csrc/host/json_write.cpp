#include "json_write.h"

#include <cmath>
#include <cstring>

namespace rin {
namespace json {

namespace {

// a floating-point value f * 2^e with a 64-bit significand
struct Fp {
    uint64_t f;
    int e;
};

Fp mul(Fp a, Fp b) {
    const unsigned __int128 p = (unsigned __int128)a.f * b.f;
    uint64_t hi = (uint64_t)(p >> 64);
    const uint64_t lo = (uint64_t)p;
    hi += lo >> 63;  // round half up
    return {hi, a.e + b.e + 64};
}

Fp normalize(Fp x) {
    while (!(x.f >> 63)) {
        x.f <<= 1;
        --x.e;
    }
    return x;
}

// 10^k for k = -300, -292, ..., 324: significand rounded to nearest, with its
// top bit set, and binary exponent (value = f * 2^e)
struct Pow10 {
    uint64_t f;
    int e, k;
};
constexpr Pow10 kPow10[] = {
    {0xAB70FE17C79AC6CAULL, -1060, -300},
    {0xFF77B1FCBEBCDC4FULL, -1034, -292},
    {0xBE5691EF416BD60CULL, -1007, -284},
    {0x8DD01FAD907FFC3CULL, -980, -276},
    {0xD3515C2831559A83ULL, -954, -268},
    {0x9D71AC8FADA6C9B5ULL, -927, -260},
    {0xEA9C227723EE8BCBULL, -901, -252},
    {0xAECC49914078536DULL, -874, -244},
    {0x823C12795DB6CE57ULL, -847, -236},
    {0xC21094364DFB5637ULL, -821, -228},
    {0x9096EA6F3848984FULL, -794, -220},
    {0xD77485CB25823AC7ULL, -768, -212},
    {0xA086CFCD97BF97F4ULL, -741, -204},
    {0xEF340A98172AACE5ULL, -715, -196},
    {0xB23867FB2A35B28EULL, -688, -188},
    {0x84C8D4DFD2C63F3BULL, -661, -180},
    {0xC5DD44271AD3CDBAULL, -635, -172},
    {0x936B9FCEBB25C996ULL, -608, -164},
    {0xDBAC6C247D62A584ULL, -582, -156},
    {0xA3AB66580D5FDAF6ULL, -555, -148},
    {0xF3E2F893DEC3F126ULL, -529, -140},
    {0xB5B5ADA8AAFF80B8ULL, -502, -132},
    {0x87625F056C7C4A8BULL, -475, -124},
    {0xC9BCFF6034C13053ULL, -449, -116},
    {0x964E858C91BA2655ULL, -422, -108},
    {0xDFF9772470297EBDULL, -396, -100},
    {0xA6DFBD9FB8E5B88FULL, -369, -92},
    {0xF8A95FCF88747D94ULL, -343, -84},
    {0xB94470938FA89BCFULL, -316, -76},
    {0x8A08F0F8BF0F156BULL, -289, -68},
    {0xCDB02555653131B6ULL, -263, -60},
    {0x993FE2C6D07B7FACULL, -236, -52},
    {0xE45C10C42A2B3B06ULL, -210, -44},
    {0xAA242499697392D3ULL, -183, -36},
    {0xFD87B5F28300CA0EULL, -157, -28},
    {0xBCE5086492111AEBULL, -130, -20},
    {0x8CBCCC096F5088CCULL, -103, -12},
    {0xD1B71758E219652CULL, -77, -4},
    {0x9C40000000000000ULL, -50, 4},
    {0xE8D4A51000000000ULL, -24, 12},
    {0xAD78EBC5AC620000ULL, 3, 20},
    {0x813F3978F8940984ULL, 30, 28},
    {0xC097CE7BC90715B3ULL, 56, 36},
    {0x8F7E32CE7BEA5C70ULL, 83, 44},
    {0xD5D238A4ABE98068ULL, 109, 52},
    {0x9F4F2726179A2245ULL, 136, 60},
    {0xED63A231D4C4FB27ULL, 162, 68},
    {0xB0DE65388CC8ADA8ULL, 189, 76},
    {0x83C7088E1AAB65DBULL, 216, 84},
    {0xC45D1DF942711D9AULL, 242, 92},
    {0x924D692CA61BE758ULL, 269, 100},
    {0xDA01EE641A708DEAULL, 295, 108},
    {0xA26DA3999AEF774AULL, 322, 116},
    {0xF209787BB47D6B85ULL, 348, 124},
    {0xB454E4A179DD1877ULL, 375, 132},
    {0x865B86925B9BC5C2ULL, 402, 140},
    {0xC83553C5C8965D3DULL, 428, 148},
    {0x952AB45CFA97A0B3ULL, 455, 156},
    {0xDE469FBD99A05FE3ULL, 481, 164},
    {0xA59BC234DB398C25ULL, 508, 172},
    {0xF6C69A72A3989F5CULL, 534, 180},
    {0xB7DCBF5354E9BECEULL, 561, 188},
    {0x88FCF317F22241E2ULL, 588, 196},
    {0xCC20CE9BD35C78A5ULL, 614, 204},
    {0x98165AF37B2153DFULL, 641, 212},
    {0xE2A0B5DC971F303AULL, 667, 220},
    {0xA8D9D1535CE3B396ULL, 694, 228},
    {0xFB9B7CD9A4A7443CULL, 720, 236},
    {0xBB764C4CA7A44410ULL, 747, 244},
    {0x8BAB8EEFB6409C1AULL, 774, 252},
    {0xD01FEF10A657842CULL, 800, 260},
    {0x9B10A4E5E9913129ULL, 827, 268},
    {0xE7109BFBA19C0C9DULL, 853, 276},
    {0xAC2820D9623BF429ULL, 880, 284},
    {0x80444B5E7AA7CF85ULL, 907, 292},
    {0xBF21E44003ACDD2DULL, 933, 300},
    {0x8E679C2F5E44FF8FULL, 960, 308},
    {0xD433179D9C8CB841ULL, 986, 316},
    {0x9E19DB92B4E31BA9ULL, 1013, 324},
};
constexpr int kPow10MinK = -300, kPow10Step = 8;
// target window of the scaled exponent (Grisu's alpha and gamma)
constexpr int kAlpha = -60, kGamma = -32;

// the cached power c = 10^-k that brings a value with binary exponent e into
// [kAlpha, kGamma]: k = ceil((kAlpha - e - 1) * log10(2)), rounded up to the
// table's step
const Pow10& cached_power(int e) {
    const int f = kAlpha - e - 1;
    const int k = (f * 78913) / (1 << 18) + (f > 0);
    const int idx = (-kPow10MinK + k + (kPow10Step - 1)) / kPow10Step;
    return kPow10[idx];
}

// move the last digit down while the number stays within the boundaries and
// gets closer to the exact value
void round_last(char* buf, int len, uint64_t dist, uint64_t delta, uint64_t rest, uint64_t ten_k) {
    while (rest < dist && delta - rest >= ten_k && (rest + ten_k < dist || dist - rest > rest + ten_k - dist)) {
        --buf[len - 1];
        rest += ten_k;
    }
}

// digits of the shortest decimal within (M-, M+) near w; all three share M+'s
// exponent, which lies in [kAlpha, kGamma]
void digits(char* buf, int& len, int& dexp, Fp Mm, Fp w, Fp Mp) {
    uint64_t delta = Mp.f - Mm.f;
    uint64_t dist = Mp.f - w.f;
    const int sh = -Mp.e;
    const uint64_t one = 1ull << sh;
    uint32_t p1 = (uint32_t)(Mp.f >> sh);
    uint64_t p2 = Mp.f & (one - 1);
    // integral digits
    uint32_t pow10 = 1;
    int n = 1;
    while (n < 10 && p1 >= pow10 * 10u) {
        pow10 *= 10;
        ++n;
    }
    while (n > 0) {
        const uint32_t d = p1 / pow10;
        p1 %= pow10;
        buf[len++] = (char)('0' + d);
        --n;
        const uint64_t rest = ((uint64_t)p1 << sh) + p2;
        if (rest <= delta) {
            dexp += n;
            round_last(buf, len, dist, delta, rest, (uint64_t)pow10 << sh);
            return;
        }
        pow10 /= 10;
    }
    // fractional digits
    int m = 0;
    while (true) {
        p2 *= 10;
        const uint64_t d = p2 >> sh;
        p2 &= one - 1;
        buf[len++] = (char)('0' + d);
        ++m;
        delta *= 10;
        dist *= 10;
        if (p2 <= delta) break;
    }
    dexp -= m;
    round_last(buf, len, dist, delta, p2, one);
}

// positive finite v -> digits and decimal exponent (v ~ digits * 10^dexp)
void shortest(double v, char* buf, int& len, int& dexp) {
    uint64_t bits;
    std::memcpy(&bits, &v, 8);
    const uint64_t F = bits & ((1ull << 52) - 1);
    const int E = (int)(bits >> 52);
    Fp x = E == 0 ? Fp{F, 1 - 1075} : Fp{F | (1ull << 52), E - 1075};
    const bool lower_closer = F == 0 && E > 1;
    Fp mp = normalize(Fp{2 * x.f + 1, x.e - 1});
    Fp mm = lower_closer ? Fp{4 * x.f - 1, x.e - 2} : Fp{2 * x.f - 1, x.e - 1};
    mm.f <<= (mm.e - mp.e);
    mm.e = mp.e;
    const Fp w = normalize(x);
    const Pow10& c = cached_power(mp.e);
    const Fp cf{c.f, c.e};
    const Fp W = mul(w, cf), Wm = mul(mm, cf), Wp = mul(mp, cf);
    len = 0;
    dexp = -c.k;
    digits(buf, len, dexp, Fp{Wm.f + 1, Wm.e}, W, Fp{Wp.f - 1, Wp.e});
}

void put_exponent(std::string& out, int e) {
    out += e < 0 ? '-' : '+';
    const int a = e < 0 ? -e : e;
    if (a < 10) {
        out += '0';
        out += (char)('0' + a);
    } else if (a < 100) {
        out += (char)('0' + a / 10);
        out += (char)('0' + a % 10);
    } else {
        out += (char)('0' + a / 100);
        out += (char)('0' + a / 10 % 10);
        out += (char)('0' + a % 10);
    }
}

}  // namespace

void put_double(std::string& out, double v) {
    if (!std::isfinite(v)) {
        out += "null";
        return;
    }
    if (std::signbit(v)) {
        out += '-';
        v = -v;
    }
    if (v == 0) {
        out += "0.0";
        return;
    }
    char d[32];
    int k = 0, dexp = 0;
    shortest(v, d, k, dexp);
    const int n = k + dexp;  // position of the decimal point
    constexpr int kMinExp = -4, kMaxExp = 15;
    if (k <= n && n <= kMaxExp) {
        out.append(d, k);
        out.append((size_t)(n - k), '0');
        out += ".0";
    } else if (0 < n && n <= kMaxExp) {
        out.append(d, n);
        out += '.';
        out.append(d + n, k - n);
    } else if (kMinExp < n && n <= 0) {
        out += "0.";
        out.append((size_t)(-n), '0');
        out.append(d, k);
    } else {
        out += d[0];
        if (k > 1) {
            out += '.';
            out.append(d + 1, k - 1);
        }
        out += 'e';
        put_exponent(out, n - 1);
    }
}

void put_uint(std::string& out, uint64_t v) {
    char b[24];
    int n = 0;
    do {
        b[n++] = (char)('0' + v % 10);
        v /= 10;
    } while (v);
    while (n) out += b[--n];
}

void put_string(std::string& out, const std::string& s) {
    static const char* hex = "0123456789abcdef";
    out += '"';
    for (unsigned char c : s) {
        switch (c) {
            case '"': out += "\\\""; break;
            case '\\': out += "\\\\"; break;
            case '\b': out += "\\b"; break;
            case '\f': out += "\\f"; break;
            case '\n': out += "\\n"; break;
            case '\r': out += "\\r"; break;
            case '\t': out += "\\t"; break;
            default:
                if (c < 0x20) {
                    out += "\\u00";
                    out += hex[c >> 4];
                    out += hex[c & 15];
                } else {
                    out += (char)c;
                }
        }
    }
    out += '"';
}

}  // namespace json
}  // namespace rin

// cz_lz4.cpp -- the LZ4 second stage of the .cz writer ("Encoder: lz4").
//
// The reference writer calls LZ4_compress (CompressionEncoders.h:230-231) from
// the LZ4 release it vendors (ThirdParty/lz4/lz4.c, 2011-2012).  A .cz file is
// only byte-compatible with the reference when every chunk is compressed by
// the same greedy parse, so this encoder makes the same choices:
//   - a 4-byte multiplicative hash (2654435761) into 2^12 u32 slots, or 2^13
//     u16 slots for inputs shorter than 65536 + 11 bytes;
//   - match search that skips ahead by (attempts++ >> 6) with attempts
//     starting at 67, matches of >= 4 bytes within 65535 bytes back (the
//     small-input table only holds 16-bit positions and checks no distance);
//   - backward extension of a match while the preceding bytes agree;
//   - after a match, position ip-2 is hashed and ip is tried at once;
//   - no match may start within the last 12 bytes, none may extend into the
//     last 5 (literals only there).
// The block format (token, literal and match length extensions, 16-bit
// little-endian offsets) is the LZ4 block format; the decoder below is a plain
// bounds-checked decoder of it.
#include <cstdint>
#include <cstring>
#include <vector>

namespace cz {

namespace {

constexpr int MINMATCH = 4;
constexpr int LASTLITERALS = 5;
constexpr int MFLIMIT = 12;  // 8 + MINMATCH
constexpr int MINLENGTH = MFLIMIT + 1;
constexpr int MAX_DISTANCE = 65535;
constexpr int SKIP = 6;
constexpr size_t SMALL_LIMIT = 65536 + (MFLIMIT - 1);

inline uint32_t rd32(const uint8_t *p)
{
    uint32_t v;
    std::memcpy(&v, p, 4);
    return v;
}
inline uint64_t rd64(const uint8_t *p)
{
    uint64_t v;
    std::memcpy(&v, p, 8);
    return v;
}
inline uint16_t rd16(const uint8_t *p)
{
    uint16_t v;
    std::memcpy(&v, p, 2);
    return v;
}

template <int LOG>
inline uint32_t hash4(const uint8_t *p)
{
    return (rd32(p) * 2654435761u) >> (32 - LOG);
}

struct Out {
    std::vector<uint8_t> &v;
    void put(uint8_t b) { v.push_back(b); }
    void lengths(size_t len)  // the 255-run extension after a 15 in the token
    {
        for (; len > 254; len -= 255) put(255);
        put((uint8_t)len);
    }
};

// One greedy LZ4 parse.  SMALL selects the 16-bit table of short inputs.
template <bool SMALL>
void compress_impl(const uint8_t *src, size_t n, std::vector<uint8_t> &dst)
{
    constexpr int LOG = SMALL ? 13 : 12;
    std::vector<uint32_t> table(size_t(1) << LOG, 0);  // positions (u16 range when SMALL)
    Out o{dst};
    size_t anchor = 0;
    const long long mflimit = (long long)n - MFLIMIT;
    const long long matchlimit = (long long)n - LASTLITERALS;
    auto far = [&](long long ref, long long ip) { return !SMALL && ref < ip - MAX_DISTANCE; };

    if (n >= (size_t)MINLENGTH) {
        long long ip = 0;
        if (!SMALL) table[hash4<LOG>(src)] = 0;  // the small-input parse leaves position 0 implicit
        ++ip;
        uint32_t fh = hash4<LOG>(src + ip);
        for (;;) {
            int attempts = (1 << SKIP) + 3;
            long long fwd = ip, ref;
            size_t token;
            bool done = false;
            do {  // find a match
                const uint32_t h = fh;
                const int step = attempts++ >> SKIP;
                ip = fwd;
                fwd = ip + step;
                if (fwd > mflimit) {
                    done = true;
                    break;
                }
                fh = hash4<LOG>(src + fwd);
                ref = table[h];
                table[h] = (uint32_t)(SMALL ? (uint16_t)ip : ip);
            } while (far(ref, ip) || rd32(src + ref) != rd32(src + ip));
            if (done) break;
            while (ip > (long long)anchor && ref > 0 && src[ip - 1] == src[ref - 1]) {  // extend backwards
                --ip;
                --ref;
            }
            {  // literals
                const size_t lit = (size_t)ip - anchor;
                token = dst.size();
                if (lit >= 15) {
                    o.put(15 << 4);
                    o.lengths(lit - 15);
                } else {
                    o.put((uint8_t)(lit << 4));
                }
                dst.insert(dst.end(), src + anchor, src + ip);
            }
            for (;;) {  // a match, then possibly another one at once
                const uint16_t off = (uint16_t)(ip - ref);
                o.put((uint8_t)off);
                o.put((uint8_t)(off >> 8));
                ip += MINMATCH;
                ref += MINMATCH;
                const long long mstart = ip;
                bool counted = false;
                while (ip < matchlimit - 7) {
                    const uint64_t diff = rd64(src + ref) ^ rd64(src + ip);
                    if (!diff) {
                        ip += 8;
                        ref += 8;
                        continue;
                    }
                    ip += __builtin_ctzll(diff) >> 3;
                    counted = true;
                    break;
                }
                if (!counted) {
                    if (ip < matchlimit - 3 && rd32(src + ref) == rd32(src + ip)) {
                        ip += 4;
                        ref += 4;
                    }
                    if (ip < matchlimit - 1 && rd16(src + ref) == rd16(src + ip)) {
                        ip += 2;
                        ref += 2;
                    }
                    if (ip < matchlimit && src[ref] == src[ip]) ++ip;
                }
                size_t ml = (size_t)(ip - mstart);
                if (ml >= 15) {
                    dst[token] += 15;
                    o.lengths(ml - 15);
                } else {
                    dst[token] += (uint8_t)ml;
                }
                if (ip > mflimit) {
                    anchor = (size_t)ip;
                    done = true;
                    break;
                }
                table[hash4<LOG>(src + ip - 2)] = (uint32_t)(SMALL ? (uint16_t)(ip - 2) : ip - 2);
                const uint32_t h = hash4<LOG>(src + ip);
                ref = table[h];
                table[h] = (uint32_t)(SMALL ? (uint16_t)ip : ip);
                if ((SMALL || ref > ip - (MAX_DISTANCE + 1)) && rd32(src + ref) == rd32(src + ip)) {
                    token = dst.size();
                    o.put(0);
                    continue;
                }
                anchor = (size_t)ip++;
                fh = hash4<LOG>(src + ip);
                break;
            }
            if (done) break;
        }
    }
    // last literals
    const size_t last = n - anchor;
    if (last >= 15) {
        o.put(15 << 4);
        o.lengths(last - 15);
    } else {
        o.put((uint8_t)(last << 4));
    }
    dst.insert(dst.end(), src + anchor, src + n);
}

}  // namespace

void lz4_compress(const uint8_t *src, size_t n, std::vector<uint8_t> &dst)
{
    dst.clear();
    dst.reserve(n + n / 255 + 16);
    if (n < SMALL_LIMIT) compress_impl<true>(src, n, dst);
    else compress_impl<false>(src, n, dst);
}

bool lz4_decompress(const uint8_t *src, size_t n, std::vector<uint8_t> &dst, size_t max_out)
{
    dst.clear();
    size_t ip = 0;
    while (ip < n) {
        const uint8_t token = src[ip++];
        size_t lit = token >> 4;
        if (lit == 15) {
            uint8_t b;
            do {
                if (ip >= n) return false;
                b = src[ip++];
                lit += b;
            } while (b == 255);
        }
        if (lit > n - ip || dst.size() + lit > max_out) return false;
        dst.insert(dst.end(), src + ip, src + ip + lit);
        ip += lit;
        if (ip == n) return true;  // the last sequence holds literals only
        if (n - ip < 2) return false;
        const size_t off = (size_t)src[ip] | ((size_t)src[ip + 1] << 8);
        ip += 2;
        size_t ml = (token & 15) + MINMATCH;
        if ((token & 15) == 15) {
            uint8_t b;
            do {
                if (ip >= n) return false;
                b = src[ip++];
                ml += b;
            } while (b == 255);
        }
        if (off == 0 || off > dst.size() || dst.size() + ml > max_out) return false;
        const size_t from = dst.size() - off;
        for (size_t k = 0; k < ml; ++k) dst.push_back(dst[from + k]);  // may overlap its own output
    }
    return n > 0 ? false : true;
}

}  // namespace cz

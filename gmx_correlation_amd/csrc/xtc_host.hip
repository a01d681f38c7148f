// XTC trajectory frames (GROMACS' compressed coordinate format) for real
// trajectory ingest without GROMACS.  The reference takes its frames from the
// GROMACS 2025 analysis framework (src/gmx_correlation.cpp:158-182); XTC is
// that framework's usual input format.  Format (XDR, big-endian words):
//
//   frame  : int magic 1995, int natoms, int step, float time,
//            float box[3][3], then the coordinate block;
//   block  : int natoms; natoms <= 9: 3*natoms floats; otherwise
//            float precision, int minint[3], int maxint[3], int smallidx,
//            int nbytes, nbytes of bit stream (padded to 4 bytes).
//
// The bit stream (most significant bit first) holds the coordinates as
// integers round(x * precision), offset by minint.  Each step writes one
// atom "large" (its three offsets in mixed radix of sizeint = maxint - minint
// + 1, or per coordinate when a size exceeds 2^24), a flag bit, and when the
// flag is set a 5-bit code (run + is_smaller + 1).  `run` (a multiple of 3,
// up to 24) coordinates of following atoms then come as "small" differences
// to the previous atom, in mixed radix of magicints[smallidx] each, offset by
// magicints[smallidx] / 2; the first small atom is stored before the large
// one (the writer swaps them: water oxygen and hydrogen).  is_smaller moves
// smallidx by -1, 0 or +1 after the step.  The magicints table, the run and
// smallidx rules are the format's; the code is this library's own.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "gmxcorr.h"

namespace gmxcorr {
int set_error(int code, const std::string& msg);
}

namespace {

// magicints[i]^3 fits in i bits (i >= 9); the table is part of the format.
const int kMagic[] = {
    0,        0,        0,       0,       0,       0,       0,       0,       0,       8,
    10,       12,       16,      20,      25,      32,      40,      50,      64,      80,
    101,      128,      161,     203,     256,     322,     406,     512,     645,     812,
    1024,     1290,     1625,    2048,    2580,    3250,    4096,    5060,    6501,    8192,
    10321,    13003,    16384,   20642,   26007,   32768,   41285,   52015,   65536,   82570,
    104031,   131072,   165140,  208063,  262144,  330280,  416127,  524287,  660561,  832255,
    1048576,  1321122,  1664510, 2097152, 2642245, 3329021, 4194304, 5284491, 6658042, 8388607,
    10568983, 13316085, 16777216};
constexpr int kFirstIdx = 9;
constexpr int kLastIdx = (int)(sizeof(kMagic) / sizeof(kMagic[0]));   // 73
constexpr int kXtcMagic = 1995;

typedef unsigned __int128 u128;

// Bits needed for values below `size` (bit length of size).
int bits_of(unsigned long long size)
{
    int b = 0;
    while (b < 64 && (size >> b) != 0) ++b;
    return b;
}

// Bit length of the product of three sizes (each below 2^24).
int bits_of3(const unsigned s[3])
{
    const u128 p = (u128)s[0] * s[1] * s[2];
    int b = 0;
    while (b < 128 && (p >> b) != 0) ++b;
    return b;
}

struct BitWriter {
    std::vector<unsigned char> bytes;
    unsigned cur = 0;   // pending bits, right-aligned
    int ncur = 0;
    void put(unsigned long long v, int nbits)   // the low nbits of v, MSB first
    {
        for (int b = nbits - 1; b >= 0; --b) {
            cur = (cur << 1) | (unsigned)((v >> b) & 1u);
            if (++ncur == 8) {
                bytes.push_back((unsigned char)cur);
                cur = 0;
                ncur = 0;
            }
        }
    }
    void flush()
    {
        if (ncur) {
            bytes.push_back((unsigned char)(cur << (8 - ncur)));
            cur = 0;
            ncur = 0;
        }
    }
};

struct BitReader {
    const unsigned char* p;
    size_t n;
    size_t pos = 0;   // bit position
    bool overrun = false;
    unsigned long long get(int nbits)
    {
        unsigned long long v = 0;
        for (int b = 0; b < nbits; ++b) {
            const size_t byte = pos >> 3;
            unsigned bit = 0;
            if (byte < n) bit = (p[byte] >> (7 - (pos & 7))) & 1u;
            else overrun = true;
            v = (v << 1) | bit;
            ++pos;
        }
        return v;
    }
};

// Three values in mixed radix (v0 * s1 + v1) * s2 + v2, written as
// little-endian bytes of nbits bits in total (the last byte short).
void put_ints3(BitWriter& w, int nbits, const unsigned s[3], const unsigned v[3])
{
    const u128 num = ((u128)v[0] * s[1] + v[1]) * s[2] + v[2];
    for (int c = 0; 8 * c < nbits; ++c) {
        const int nb = std::min(8, nbits - 8 * c);
        w.put((unsigned long long)((num >> (8 * c)) & 0xffu), nb);
    }
}

void get_ints3(BitReader& r, int nbits, const unsigned s[3], int v[3])
{
    u128 num = 0;
    for (int c = 0; 8 * c < nbits; ++c) {
        const int nb = std::min(8, nbits - 8 * c);
        num |= (u128)r.get(nb) << (8 * c);
    }
    v[2] = (int)(num % s[2]);
    num /= s[2];
    v[1] = (int)(num % s[1]);
    num /= s[1];
    v[0] = (int)num;
}

void put_u32(std::vector<unsigned char>& o, uint32_t v)
{
    o.push_back((unsigned char)(v >> 24));
    o.push_back((unsigned char)(v >> 16));
    o.push_back((unsigned char)(v >> 8));
    o.push_back((unsigned char)v);
}
void put_i32(std::vector<unsigned char>& o, int v) { put_u32(o, (uint32_t)v); }
void put_f32(std::vector<unsigned char>& o, float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    put_u32(o, u);
}

bool read_u32(FILE* fp, uint32_t& v)
{
    unsigned char b[4];
    if (std::fread(b, 1, 4, fp) != 4) return false;
    v = ((uint32_t)b[0] << 24) | ((uint32_t)b[1] << 16) | ((uint32_t)b[2] << 8) | b[3];
    return true;
}
bool read_i32(FILE* fp, int& v)
{
    uint32_t u;
    if (!read_u32(fp, u)) return false;
    v = (int)u;
    return true;
}
bool read_f32(FILE* fp, float& f)
{
    uint32_t u;
    if (!read_u32(fp, u)) return false;
    std::memcpy(&f, &u, 4);
    return true;
}

int fail(int code, const std::string& msg) { return gmxcorr::set_error(code, msg); }

// Coordinate block of one frame (after its header).
int encode_coords(std::vector<unsigned char>& o, int natoms, const float* x, float precision)
{
    put_i32(o, natoms);
    if (natoms <= 9) {
        for (int t = 0; t < 3 * natoms; ++t) put_f32(o, x[t]);
        return GMXCORR_OK;
    }
    if (!(precision > 0.0f)) return fail(GMXCORR_EINVAL, "xtc: precision must be positive");
    std::vector<int> ip((size_t)3 * natoms);
    int minint[3] = {INT_MAX, INT_MAX, INT_MAX}, maxint[3] = {INT_MIN, INT_MIN, INT_MIN};
    long long mindiff = LLONG_MAX;
    for (int a = 0; a < natoms; ++a) {
        for (int d = 0; d < 3; ++d) {
            const float lf = x[3 * a + d] * precision;
            if (!(std::fabs(lf) < (float)(INT_MAX - 2)))
                return fail(GMXCORR_EINVAL, "xtc: coordinate too large for the precision");
            const int li = (int)(lf >= 0.0f ? lf + 0.5f : lf - 0.5f);
            ip[3 * a + d] = li;
            minint[d] = std::min(minint[d], li);
            maxint[d] = std::max(maxint[d], li);
        }
        if (a > 0) {
            long long diff = 0;
            for (int d = 0; d < 3; ++d)
                diff += std::llabs((long long)ip[3 * a + d] - ip[3 * (a - 1) + d]);
            mindiff = std::min(mindiff, diff);
        }
    }
    put_f32(o, precision);
    for (int d = 0; d < 3; ++d) put_i32(o, minint[d]);
    for (int d = 0; d < 3; ++d) put_i32(o, maxint[d]);
    unsigned sizeint[3];
    int bitsizeint[3] = {0, 0, 0}, bitsize = 0;
    bool large = false;
    for (int d = 0; d < 3; ++d) {
        const long long s = (long long)maxint[d] - minint[d] + 1;
        if (s > 0xffffff) large = true;
        sizeint[d] = (unsigned)std::min<long long>(s, 0xffffffffLL);
    }
    if (large)
        for (int d = 0; d < 3; ++d) bitsizeint[d] = bits_of(sizeint[d]);
    else
        bitsize = bits_of3(sizeint);
    int smallidx = kFirstIdx;
    while (smallidx < kLastIdx - 1 && kMagic[smallidx] < mindiff) ++smallidx;
    put_i32(o, smallidx);
    const int maxidx = std::min(kLastIdx - 1, smallidx + 8);
    const int minidx = maxidx - 8;
    int smaller = kMagic[std::max(kFirstIdx, smallidx - 1)] / 2;
    int smallnum = kMagic[smallidx] / 2;
    unsigned sizesmall[3] = {(unsigned)kMagic[smallidx], (unsigned)kMagic[smallidx],
                             (unsigned)kMagic[smallidx]};
    const int larger = kMagic[maxidx] / 2;
    BitWriter w;
    int i = 0, prevrun = -1;
    int prev[3] = {0, 0, 0};
    auto absd = [](int u, int v) { return std::llabs((long long)u - v); };
    while (i < natoms) {
        int* cur = &ip[3 * i];
        int is_smaller;
        if (smallidx < maxidx && i >= 1 && absd(cur[0], prev[0]) < larger &&
            absd(cur[1], prev[1]) < larger && absd(cur[2], prev[2]) < larger)
            is_smaller = 1;
        else if (smallidx > minidx)
            is_smaller = -1;
        else
            is_smaller = 0;
        bool is_small = false;
        if (i + 1 < natoms && absd(cur[0], cur[3]) < smallnum && absd(cur[1], cur[4]) < smallnum &&
            absd(cur[2], cur[5]) < smallnum) {
            // the second atom goes first, as a small step from this one
            for (int d = 0; d < 3; ++d) std::swap(cur[d], cur[3 + d]);
            is_small = true;
        }
        unsigned t3[3];
        for (int d = 0; d < 3; ++d) t3[d] = (unsigned)(cur[d] - minint[d]);
        if (large)
            for (int d = 0; d < 3; ++d) w.put(t3[d], bitsizeint[d]);
        else
            put_ints3(w, bitsize, sizeint, t3);
        for (int d = 0; d < 3; ++d) prev[d] = cur[d];
        ++i;
        int run = 0;
        unsigned tmp[24];
        if (!is_small && is_smaller == -1) is_smaller = 0;
        while (is_small && run < 24) {
            int* c2 = &ip[3 * i];
            if (is_smaller == -1) {
                long long d2 = 0;
                for (int d = 0; d < 3; ++d) d2 += ((long long)c2[d] - prev[d]) * ((long long)c2[d] - prev[d]);
                if (d2 >= (long long)smaller * smaller) is_smaller = 0;
            }
            for (int d = 0; d < 3; ++d) tmp[run++] = (unsigned)(c2[d] - prev[d] + smallnum);
            for (int d = 0; d < 3; ++d) prev[d] = c2[d];
            ++i;
            is_small = false;
            if (i < natoms) {
                const int* c3 = &ip[3 * i];
                if (absd(c3[0], prev[0]) < smallnum && absd(c3[1], prev[1]) < smallnum &&
                    absd(c3[2], prev[2]) < smallnum)
                    is_small = true;
            }
        }
        if (run != prevrun || is_smaller != 0) {
            prevrun = run;
            w.put(1, 1);
            w.put((unsigned)(run + is_smaller + 1), 5);
        } else {
            w.put(0, 1);
        }
        for (int t = 0; t < run; t += 3) put_ints3(w, smallidx, sizesmall, &tmp[t]);
        if (is_smaller != 0) {
            smallidx += is_smaller;
            if (is_smaller < 0) {
                smallnum = smaller;
                smaller = smallidx > kFirstIdx ? kMagic[smallidx - 1] / 2 : 0;
            } else {
                smaller = smallnum;
                smallnum = kMagic[smallidx] / 2;
            }
            sizesmall[0] = sizesmall[1] = sizesmall[2] = (unsigned)kMagic[smallidx];
        }
    }
    w.flush();
    put_i32(o, (int)w.bytes.size());
    o.insert(o.end(), w.bytes.begin(), w.bytes.end());
    while (o.size() % 4) o.push_back(0);
    return GMXCORR_OK;
}

int decode_coords(FILE* fp, int natoms, float* x, float* precision_out)
{
    int lsize = 0;
    if (!read_i32(fp, lsize)) return fail(GMXCORR_EINVAL, "xtc: truncated coordinate block");
    if (lsize != natoms) return fail(GMXCORR_EINVAL, "xtc: atom counts of header and block differ");
    if (natoms <= 9) {
        for (int t = 0; t < 3 * natoms; ++t)
            if (!read_f32(fp, x[t])) return fail(GMXCORR_EINVAL, "xtc: truncated coordinates");
        if (precision_out) *precision_out = 0.0f;
        return GMXCORR_OK;
    }
    float precision = 0.0f;
    int minint[3], maxint[3], smallidx = 0, nbytes = 0;
    bool ok = read_f32(fp, precision);
    for (int d = 0; d < 3; ++d) ok = ok && read_i32(fp, minint[d]);
    for (int d = 0; d < 3; ++d) ok = ok && read_i32(fp, maxint[d]);
    ok = ok && read_i32(fp, smallidx) && read_i32(fp, nbytes);
    if (!ok) return fail(GMXCORR_EINVAL, "xtc: truncated compressed header");
    if (!(precision > 0.0f) || smallidx < kFirstIdx || smallidx >= kLastIdx || nbytes < 0)
        return fail(GMXCORR_EINVAL, "xtc: corrupt compressed header");
    std::vector<unsigned char> buf(((size_t)nbytes + 3) / 4 * 4);
    if (std::fread(buf.data(), 1, buf.size(), fp) != buf.size())
        return fail(GMXCORR_EINVAL, "xtc: truncated bit stream");
    unsigned sizeint[3];
    int bitsizeint[3] = {0, 0, 0}, bitsize = 0;
    bool large = false;
    for (int d = 0; d < 3; ++d) {
        const long long s = (long long)maxint[d] - minint[d] + 1;
        if (s <= 0) return fail(GMXCORR_EINVAL, "xtc: corrupt coordinate range");
        if (s > 0xffffff) large = true;
        sizeint[d] = (unsigned)std::min<long long>(s, 0xffffffffLL);
    }
    if (large)
        for (int d = 0; d < 3; ++d) bitsizeint[d] = bits_of(sizeint[d]);
    else
        bitsize = bits_of3(sizeint);
    int smaller = kMagic[std::max(kFirstIdx, smallidx - 1)] / 2;
    int smallnum = kMagic[smallidx] / 2;
    unsigned sizesmall[3] = {(unsigned)kMagic[smallidx], (unsigned)kMagic[smallidx],
                             (unsigned)kMagic[smallidx]};
    const float inv = 1.0f / precision;
    BitReader r{buf.data(), (size_t)nbytes};
    int i = 0, run = 0;
    float* out = x;
    while (i < natoms) {
        int cur[3], prev[3];
        if (large)
            for (int d = 0; d < 3; ++d) cur[d] = (int)r.get(bitsizeint[d]);
        else
            get_ints3(r, bitsize, sizeint, cur);
        ++i;
        for (int d = 0; d < 3; ++d) {
            cur[d] += minint[d];
            prev[d] = cur[d];
        }
        const unsigned flag = (unsigned)r.get(1);
        int is_smaller = 0;
        if (flag == 1) {
            run = (int)r.get(5);
            is_smaller = run % 3;
            run -= is_smaller;
            --is_smaller;
        }
        if (run > 0) {
            if (i - 1 + run / 3 >= natoms)
                return fail(GMXCORR_EINVAL, "xtc: run beyond the atom count");
            for (int t = 0; t < run; t += 3) {
                int sm[3];
                get_ints3(r, smallidx, sizesmall, sm);
                ++i;
                for (int d = 0; d < 3; ++d) sm[d] += prev[d] - smallnum;
                if (t == 0) {
                    // stored first: the atom written before the large one
                    for (int d = 0; d < 3; ++d) std::swap(sm[d], prev[d]);
                    for (int d = 0; d < 3; ++d) *out++ = prev[d] * inv;
                } else {
                    for (int d = 0; d < 3; ++d) prev[d] = sm[d];
                }
                for (int d = 0; d < 3; ++d) *out++ = sm[d] * inv;
            }
        } else {
            for (int d = 0; d < 3; ++d) *out++ = cur[d] * inv;
        }
        smallidx += is_smaller;
        if (smallidx < kFirstIdx || smallidx >= kLastIdx)
            return fail(GMXCORR_EINVAL, "xtc: corrupt small-step size");
        if (is_smaller < 0) {
            smallnum = smaller;
            smaller = smallidx > kFirstIdx ? kMagic[smallidx - 1] / 2 : 0;
        } else if (is_smaller > 0) {
            smaller = smallnum;
            smallnum = kMagic[smallidx] / 2;
        }
        sizesmall[0] = sizesmall[1] = sizesmall[2] = (unsigned)kMagic[smallidx];
    }
    if (r.overrun) return fail(GMXCORR_EINVAL, "xtc: bit stream shorter than its coordinates");
    if (precision_out) *precision_out = precision;
    return GMXCORR_OK;
}

}  // namespace

struct gmxcorr_xtc {
    FILE* fp = nullptr;
    bool writing = false;
};

extern "C" {

int gmxcorr_xtc_open(const char* path, int write, gmxcorr_xtc** out)
{
    if (!path || !out) return fail(GMXCORR_EINVAL, "xtc_open: null argument");
    *out = nullptr;
    FILE* fp = std::fopen(path, write ? "wb" : "rb");
    if (!fp) return fail(GMXCORR_EINVAL, std::string("xtc_open: cannot open ") + path);
    auto* f = new gmxcorr_xtc();
    f->fp = fp;
    f->writing = write != 0;
    *out = f;
    return GMXCORR_OK;
}

int gmxcorr_xtc_write(gmxcorr_xtc* f, int natoms, int step, float time, const float* box,
                      const float* x, float precision)
{
    if (!f || !f->writing || !x || natoms < 1) return fail(GMXCORR_EINVAL, "xtc_write: bad argument");
    std::vector<unsigned char> o;
    put_i32(o, kXtcMagic);
    put_i32(o, natoms);
    put_i32(o, step);
    put_f32(o, time);
    for (int t = 0; t < 9; ++t) put_f32(o, box ? box[t] : 0.0f);
    if (int rc = encode_coords(o, natoms, x, precision)) return rc;
    if (std::fwrite(o.data(), 1, o.size(), f->fp) != o.size())
        return fail(GMXCORR_EINVAL, "xtc_write: write failed");
    return GMXCORR_OK;
}

int gmxcorr_xtc_read(gmxcorr_xtc* f, int* natoms, int* step, float* time, float* box, float* x,
                     int max_atoms, float* precision)
{
    if (!f || f->writing || !natoms) return fail(GMXCORR_EINVAL, "xtc_read: bad argument");
    const off_t start = ftello(f->fp);
    int magic = 0;
    if (!read_i32(f->fp, magic)) return GMXCORR_EOF;
    if (magic != kXtcMagic) return fail(GMXCORR_EINVAL, "xtc_read: not an XTC frame (magic)");
    int n = 0, st = 0;
    float t = 0.0f, b[9];
    bool ok = read_i32(f->fp, n) && read_i32(f->fp, st) && read_f32(f->fp, t);
    for (int q = 0; q < 9; ++q) ok = ok && read_f32(f->fp, b[q]);
    if (!ok || n < 1) return fail(GMXCORR_EINVAL, "xtc_read: truncated frame header");
    *natoms = n;
    if (step) *step = st;
    if (time) *time = t;
    if (box) std::memcpy(box, b, sizeof(b));
    if (!x || max_atoms < n) {
        // the caller learns natoms; the same frame is read again next call
        (void)fseeko(f->fp, start, SEEK_SET);
        return fail(GMXCORR_EINVAL, "xtc_read: coordinate buffer too small");
    }
    return decode_coords(f->fp, n, x, precision);
}

void gmxcorr_xtc_close(gmxcorr_xtc* f)
{
    if (!f) return;
    if (f->fp) std::fclose(f->fp);
    delete f;
}

}  // extern "C"

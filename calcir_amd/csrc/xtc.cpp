// xtc.cpp -- GROMACS XTC trajectory reader and writer (host C++), the frame
// source the reference gets from xdrfile (read_xtc_natoms, read_xtc,
// read_xtc_n_frames, xdr_seek; calcIR.cu:212-219, 341-395, calcIR.h:16-18).
//
// A frame is XDR (big-endian) data:
//   int magic (1995), int natoms, int step, float time, float box[3][3],
//   then the coordinates: int natoms; natoms <= 9: 3*natoms floats; else
//   float precision, int minint[3], int maxint[3], int smallidx, int nbytes
//   and nbytes of bit-packed integers (padded to 4 bytes).
// The bit packing is the published xdr3dfcoord scheme: coordinates rounded to
// integers at the given precision, offsets from minint packed as one mixed-
// radix number per atom, and runs of atoms close to their predecessor coded
// as small differences whose radix (magicints[smallidx]) adapts along the
// frame; the first two atoms of a run are exchanged, which favours water.
// The magicints table is part of the format.

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "calcir.h"
#include "common.hpp"

namespace cir {
namespace {

constexpr int kXtcMagic = 1995;
constexpr int kFirstIdx = 9;
// radices of the small-difference code (format constant)
constexpr int kMagicInts[] = {
    0,        0,        0,       0,       0,       0,       0,       0,       0,       8,       10,      12,
    16,       20,       25,      32,      40,      50,      64,      80,      101,     128,     161,     203,
    256,      322,      406,     512,     645,     812,     1024,    1290,    1625,    2048,    2580,    3250,
    4096,     5060,     6501,    8192,    10321,   13003,   16384,   20642,   26007,   32768,   41285,   52015,
    65536,    82570,    104031,  131072,  165140,  208063,  262144,  330280,  416127,  524287,  660561,  832255,
    1048576,  1321122,  1664510, 2097152, 2642245, 3329021, 4194304, 5284491, 6658042, 8388607, 10568983,
    13316085, 16777216};
constexpr int kLastIdx = int(sizeof(kMagicInts) / sizeof(kMagicInts[0]));

// ---- XDR primitives -------------------------------------------------------
struct XdrIn {
    FILE* f;
    bool ok = true;
    uint32_t u32() {
        unsigned char b[4];
        if (std::fread(b, 1, 4, f) != 4) {
            ok = false;
            return 0;
        }
        return (uint32_t(b[0]) << 24) | (uint32_t(b[1]) << 16) | (uint32_t(b[2]) << 8) | uint32_t(b[3]);
    }
    int i32() { return int(u32()); }
    float f32() {
        const uint32_t u = u32();
        float x;
        std::memcpy(&x, &u, 4);
        return x;
    }
    bool bytes(unsigned char* p, size_t n) {
        if (n && std::fread(p, 1, n, f) != n) return ok = false;
        const size_t pad = (4 - n % 4) % 4;
        unsigned char z[4];
        if (pad && std::fread(z, 1, pad, f) != pad) return ok = false;
        return true;
    }
};

struct XdrOut {
    std::vector<unsigned char> b;
    void u32(uint32_t u) {
        b.push_back((u >> 24) & 0xff);
        b.push_back((u >> 16) & 0xff);
        b.push_back((u >> 8) & 0xff);
        b.push_back(u & 0xff);
    }
    void i32(int v) { u32(uint32_t(v)); }
    void f32(float x) {
        uint32_t u;
        std::memcpy(&u, &x, 4);
        u32(u);
    }
    void bytes(const unsigned char* p, size_t n) {
        b.insert(b.end(), p, p + n);
        for (size_t q = n; q % 4; ++q) b.push_back(0);
    }
};

// ---- bit stream -------------------------------------------------------------
struct BitReader {
    const unsigned char* p;
    size_t n, cnt = 0;
    unsigned int lastbits = 0, lastbyte = 0;
    bool ok = true;
    unsigned int next_byte() {
        if (cnt >= n) {
            ok = false;
            return 0;
        }
        return p[cnt++];
    }
    int bits(int nbits) {
        const unsigned int mask = (nbits >= 32) ? 0xffffffffu : ((1u << nbits) - 1);
        unsigned int num = 0;
        while (nbits >= 8) {
            lastbyte = (lastbyte << 8) | next_byte();
            num |= (lastbyte >> lastbits) << (nbits - 8);
            nbits -= 8;
        }
        if (nbits > 0) {
            if (int(lastbits) < nbits) {
                lastbits += 8;
                lastbyte = (lastbyte << 8) | next_byte();
            }
            lastbits -= nbits;
            num |= (lastbyte >> lastbits) & ((1u << nbits) - 1);
        }
        return int(num & mask);
    }
    // three integers packed as one mixed-radix number of nbits bits
    void ints3(int nbits, const unsigned int sizes[3], int out[3]) {
        int bytes[32] = {0};
        int nb = 0;
        while (nbits > 8) {
            bytes[nb++] = bits(8);
            nbits -= 8;
        }
        if (nbits > 0) bytes[nb++] = bits(nbits);
        for (int i = 2; i > 0; --i) {
            unsigned int num = 0;
            for (int j = nb - 1; j >= 0; --j) {
                num = (num << 8) | unsigned(bytes[j]);
                const unsigned int q = num / sizes[i];
                bytes[j] = int(q);
                num -= q * sizes[i];
            }
            out[i] = int(num);
        }
        out[0] = bytes[0] | (bytes[1] << 8) | (bytes[2] << 16) | (bytes[3] << 24);
    }
};

struct BitWriter {
    std::vector<unsigned char> buf;
    unsigned int lastbits = 0, lastbyte = 0;
    void bits(int nbits, unsigned int num) {
        while (nbits >= 8) {
            lastbyte = (lastbyte << 8) | ((num >> (nbits - 8)) & 0xff);
            buf.push_back((unsigned char)(lastbyte >> lastbits));
            nbits -= 8;
        }
        if (nbits > 0) {
            lastbyte = (lastbyte << nbits) | (num & ((1u << nbits) - 1));
            lastbits += nbits;
            if (lastbits >= 8) {
                lastbits -= 8;
                buf.push_back((unsigned char)(lastbyte >> lastbits));
            }
        }
    }
    void flush() {
        if (lastbits > 0) buf.push_back((unsigned char)(lastbyte << (8 - lastbits)));
        lastbits = 0;
    }
    void ints3(int nbits, const unsigned int sizes[3], const unsigned int nums[3]) {
        unsigned int bytes[32];
        int nb = 0;
        unsigned int tmp = nums[0];
        do {
            bytes[nb++] = tmp & 0xff;
            tmp >>= 8;
        } while (tmp != 0);
        for (int i = 1; i < 3; ++i) {
            tmp = nums[i];
            int bc = 0;
            for (; bc < nb; ++bc) {
                tmp = bytes[bc] * sizes[i] + tmp;
                bytes[bc] = tmp & 0xff;
                tmp >>= 8;
            }
            while (tmp != 0) {
                bytes[bc++] = tmp & 0xff;
                tmp >>= 8;
            }
            nb = bc;
        }
        if (nbits >= nb * 8) {
            for (int i = 0; i < nb; ++i) bits(8, bytes[i]);
            bits(nbits - nb * 8, 0);
        } else {
            int i = 0;
            for (; i < nb - 1; ++i) bits(8, bytes[i]);
            bits(nbits - (nb - 1) * 8, bytes[i]);
        }
    }
};

int size_of_int(unsigned int size) {
    unsigned int num = 1;
    int nbits = 0;
    while (size >= num && nbits < 32) {
        ++nbits;
        num <<= 1;
    }
    return nbits;
}

// bits needed for the mixed-radix number of three integers below sizes[]
int size_of_ints3(const unsigned int sizes[3]) {
    unsigned int bytes[32];
    int nb = 1;
    bytes[0] = 1;
    for (int i = 0; i < 3; ++i) {
        unsigned int tmp = 0;
        int bc = 0;
        for (; bc < nb; ++bc) {
            tmp = bytes[bc] * sizes[i] + tmp;
            bytes[bc] = tmp & 0xff;
            tmp >>= 8;
        }
        while (tmp != 0) {
            bytes[bc++] = tmp & 0xff;
            tmp >>= 8;
        }
        nb = bc;
    }
    unsigned int num = 1;
    int nbits = 0;
    --nb;
    while (bytes[nb] >= num) {
        ++nbits;
        num *= 2;
    }
    return nbits + nb * 8;
}

// The compressed block of natoms atoms holds at most 3 x 32 bits of an atom's
// full coordinates and 6 flag bits per atom (about 13 bytes; the format's
// writers size their buffer at 14.4): a larger byte count is a corrupt frame.
bool coord_bytes_ok(int nbytes, int natoms) { return nbytes >= 0 && nbytes <= 16LL * natoms + 64; }

#pragma clang fp contract(off)
float round_half_away(float x, float precision) {
    float lf = x * precision;
    lf += (lf >= 0.0f) ? 0.5f : -0.5f;
    return lf;
}

// ---- frame coordinates --------------------------------------------------------
bool read_coords(XdrIn& x, int natoms, float* out) {
    const int lsize = x.i32();
    if (!x.ok || lsize != natoms) return false;
    if (natoms <= 9) {
        for (int q = 0; q < 3 * natoms; ++q) out[q] = x.f32();
        return x.ok;
    }
    const float precision = x.f32();
    int minint[3], maxint[3];
    for (int k = 0; k < 3; ++k) minint[k] = x.i32();
    for (int k = 0; k < 3; ++k) maxint[k] = x.i32();
    if (!x.ok || !(precision > 0.0f)) return false;
    unsigned int sizeint[3];
    int bitsizeint[3] = {0, 0, 0}, bitsize = 0;
    for (int k = 0; k < 3; ++k) {
        // the writer's range limit; an empty or wider range would give a zero
        // or wrapped radix
        if (maxint[k] < minint[k] || (long long)maxint[k] - minint[k] >= 0x7fffffffLL) return false;
        sizeint[k] = unsigned(maxint[k] - minint[k]) + 1;
    }
    if ((sizeint[0] | sizeint[1] | sizeint[2]) > 0xffffff) {
        for (int k = 0; k < 3; ++k) bitsizeint[k] = size_of_int(sizeint[k]);
    } else {
        bitsize = size_of_ints3(sizeint);
    }
    int smallidx = x.i32();
    const int nbytes = x.i32();
    if (!x.ok || smallidx < kFirstIdx || smallidx >= kLastIdx || !coord_bytes_ok(nbytes, natoms)) return false;
    std::vector<unsigned char> data(size_t(nbytes) + 8, 0);
    if (!x.bytes(data.data(), size_t(nbytes))) return false;
    BitReader br{data.data(), size_t(nbytes)};
    int smaller = kMagicInts[std::max(kFirstIdx, smallidx - 1)] / 2;
    int smallnum = kMagicInts[smallidx] / 2;
    unsigned int sizesmall[3] = {unsigned(kMagicInts[smallidx]), unsigned(kMagicInts[smallidx]),
                                 unsigned(kMagicInts[smallidx])};
    const float inv = 1.0f / precision;
    int i = 0, run = 0;
    float* o = out;
    while (i < natoms) {
        int thiscoord[3], prevcoord[3];
        if (bitsize == 0) {
            for (int k = 0; k < 3; ++k) thiscoord[k] = br.bits(bitsizeint[k]);
        } else {
            br.ints3(bitsize, sizeint, thiscoord);
        }
        ++i;
        for (int k = 0; k < 3; ++k) thiscoord[k] += minint[k];
        for (int k = 0; k < 3; ++k) prevcoord[k] = thiscoord[k];
        int is_smaller = 0;
        if (br.bits(1) == 1) {
            run = br.bits(5);
            is_smaller = run % 3;
            run -= is_smaller;
            --is_smaller;
        }
        if (run > 0) {
            if (i + run / 3 > natoms) return false;  // run past the end of the frame
            for (int k = 0; k < run; k += 3) {
                br.ints3(smallidx, sizesmall, thiscoord);
                ++i;
                for (int c = 0; c < 3; ++c) thiscoord[c] += prevcoord[c] - smallnum;
                if (k == 0) {
                    // the writer exchanged the first two atoms of the run
                    for (int c = 0; c < 3; ++c) std::swap(thiscoord[c], prevcoord[c]);
                    for (int c = 0; c < 3; ++c) *o++ = float(prevcoord[c]) * inv;
                } else {
                    for (int c = 0; c < 3; ++c) prevcoord[c] = thiscoord[c];
                }
                for (int c = 0; c < 3; ++c) *o++ = float(thiscoord[c]) * inv;
            }
        } else {
            for (int c = 0; c < 3; ++c) *o++ = float(thiscoord[c]) * inv;
        }
        smallidx += is_smaller;
        if (smallidx < kFirstIdx || smallidx >= kLastIdx) return false;
        if (is_smaller < 0) {
            smallnum = smaller;
            smaller = (smallidx > kFirstIdx) ? kMagicInts[smallidx - 1] / 2 : 0;
        } else if (is_smaller > 0) {
            smaller = smallnum;
            smallnum = kMagicInts[smallidx] / 2;
        }
        sizesmall[0] = sizesmall[1] = sizesmall[2] = unsigned(kMagicInts[smallidx]);
    }
    return br.ok && (o - out) == 3 * natoms;
}

bool write_coords(XdrOut& x, int natoms, const float* fp, float precision) {
    x.i32(natoms);
    if (natoms <= 9) {
        for (int q = 0; q < 3 * natoms; ++q) x.f32(fp[q]);
        return true;
    }
    x.f32(precision);
    std::vector<int> ip(3 * size_t(natoms) + 3);
    int minint[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, maxint[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
    int mindiff = INT32_MAX;
    const float maxabs = float(INT32_MAX) - 256.0f;
    for (int a = 0; a < natoms; ++a) {
        for (int k = 0; k < 3; ++k) {
            // rounded in FP32, as the format's writers do: x*precision is
            // one float product, then +-0.5 and truncation
            const float lf = round_half_away(fp[3 * a + k], precision);
            if (!(std::fabs(lf) <= maxabs)) return false;
            const int li = int(lf);
            ip[3 * a + k] = li;
            minint[k] = std::min(minint[k], li);
            maxint[k] = std::max(maxint[k], li);
        }
        if (a > 0) {
            const long long diff = std::llabs((long long)ip[3 * a] - ip[3 * a - 3]) +
                                   std::llabs((long long)ip[3 * a + 1] - ip[3 * a - 2]) +
                                   std::llabs((long long)ip[3 * a + 2] - ip[3 * a - 1]);
            if (diff < mindiff) mindiff = int(std::min<long long>(diff, INT32_MAX));
        }
    }
    for (int k = 0; k < 3; ++k) x.i32(minint[k]);
    for (int k = 0; k < 3; ++k) x.i32(maxint[k]);
    unsigned int sizeint[3];
    int bitsizeint[3] = {0, 0, 0}, bitsize = 0;
    for (int k = 0; k < 3; ++k) {
        if ((long long)maxint[k] - minint[k] >= 0x7fffffffLL) return false;
        sizeint[k] = unsigned(maxint[k] - minint[k]) + 1;
    }
    if ((sizeint[0] | sizeint[1] | sizeint[2]) > 0xffffff) {
        for (int k = 0; k < 3; ++k) bitsizeint[k] = size_of_int(sizeint[k]);
    } else {
        bitsize = size_of_ints3(sizeint);
    }
    int smallidx = kFirstIdx;
    while (smallidx < kLastIdx - 1 && kMagicInts[smallidx] < mindiff) ++smallidx;
    x.i32(smallidx);
    const int maxidx = std::min(kLastIdx - 1, smallidx + 8);
    const int minidx = maxidx - 8;
    int smaller = kMagicInts[std::max(kFirstIdx, smallidx - 1)] / 2;
    int smallnum = kMagicInts[smallidx] / 2;
    unsigned int sizesmall[3] = {unsigned(kMagicInts[smallidx]), unsigned(kMagicInts[smallidx]),
                                 unsigned(kMagicInts[smallidx])};
    const int larger = kMagicInts[maxidx] / 2;
    BitWriter bw;
    int i = 0, prevrun = -1;
    int prevcoord[3] = {0, 0, 0};
    auto absdiff = [](int a, int b) { return std::llabs((long long)a - b); };
    while (i < natoms) {
        int is_small = 0, is_smaller;
        int* thiscoord = &ip[3 * size_t(i)];
        if (smallidx < maxidx && i >= 1 && absdiff(thiscoord[0], prevcoord[0]) < larger &&
            absdiff(thiscoord[1], prevcoord[1]) < larger && absdiff(thiscoord[2], prevcoord[2]) < larger)
            is_smaller = 1;
        else if (smallidx > minidx)
            is_smaller = -1;
        else
            is_smaller = 0;
        if (i + 1 < natoms && absdiff(thiscoord[0], thiscoord[3]) < smallnum &&
            absdiff(thiscoord[1], thiscoord[4]) < smallnum && absdiff(thiscoord[2], thiscoord[5]) < smallnum) {
            // exchange the first two atoms of the run (water: H then O)
            for (int c = 0; c < 3; ++c) std::swap(thiscoord[c], thiscoord[3 + c]);
            is_small = 1;
        }
        unsigned int tmp[3];
        for (int c = 0; c < 3; ++c) tmp[c] = unsigned(thiscoord[c] - minint[c]);
        if (bitsize == 0) {
            for (int c = 0; c < 3; ++c) bw.bits(bitsizeint[c], tmp[c]);
        } else {
            bw.ints3(bitsize, sizeint, tmp);
        }
        for (int c = 0; c < 3; ++c) prevcoord[c] = thiscoord[c];
        thiscoord += 3;
        ++i;
        int run = 0;
        unsigned int runv[8 * 3];
        if (is_small == 0 && is_smaller == -1) is_smaller = 0;
        while (is_small && run < 8 * 3) {
            if (is_smaller == -1) {
                const long long d0 = (long long)thiscoord[0] - prevcoord[0];
                const long long d1 = (long long)thiscoord[1] - prevcoord[1];
                const long long d2 = (long long)thiscoord[2] - prevcoord[2];
                if (d0 * d0 + d1 * d1 + d2 * d2 >= (long long)smaller * smaller) is_smaller = 0;
            }
            for (int c = 0; c < 3; ++c) runv[run++] = unsigned(thiscoord[c] - prevcoord[c] + smallnum);
            for (int c = 0; c < 3; ++c) prevcoord[c] = thiscoord[c];
            ++i;
            thiscoord += 3;
            is_small = 0;
            if (i < natoms && absdiff(thiscoord[0], prevcoord[0]) < smallnum &&
                absdiff(thiscoord[1], prevcoord[1]) < smallnum && absdiff(thiscoord[2], prevcoord[2]) < smallnum)
                is_small = 1;
        }
        if (run != prevrun || is_smaller != 0) {
            prevrun = run;
            bw.bits(1, 1);
            bw.bits(5, unsigned(run + is_smaller + 1));
        } else {
            bw.bits(1, 0);
        }
        for (int k = 0; k < run; k += 3) bw.ints3(smallidx, sizesmall, &runv[k]);
        if (is_smaller != 0) {
            smallidx += is_smaller;
            if (is_smaller < 0) {
                smallnum = smaller;
                smaller = (smallidx > kFirstIdx) ? kMagicInts[smallidx - 1] / 2 : 0;
            } else {
                smaller = smallnum;
                smallnum = kMagicInts[smallidx] / 2;
            }
            sizesmall[0] = sizesmall[1] = sizesmall[2] = unsigned(kMagicInts[smallidx]);
        }
    }
    bw.flush();
    x.i32(int(bw.buf.size()));
    x.bytes(bw.buf.data(), bw.buf.size());
    return true;
}

// Reads one frame header; returns false at a clean end of file.
bool read_header(XdrIn& x, int* natoms, int* step, float* time, float box[9]) {
    const int magic = x.i32();
    if (!x.ok) return false;
    if (magic != kXtcMagic) {
        x.ok = false;
        return false;
    }
    *natoms = x.i32();
    *step = x.i32();
    *time = x.f32();
    for (int q = 0; q < 9; ++q) box[q] = x.f32();
    if (x.ok && *natoms < 0) x.ok = false;  // corrupt header
    return x.ok;
}

// Skips one frame's coordinate block: its header gives the byte count of the
// compressed data, so nothing is decoded.
bool skip_coords(XdrIn& x, int natoms) {
    const int lsize = x.i32();
    if (!x.ok || lsize != natoms) return false;
    if (natoms <= 9) return std::fseek(x.f, 12L * natoms, SEEK_CUR) == 0;
    for (int k = 0; k < 8; ++k) (void)x.i32();  // precision, minint[3], maxint[3], smallidx
    const int nbytes = x.i32();
    if (!x.ok || !coord_bytes_ok(nbytes, natoms)) return false;
    return std::fseek(x.f, (long(nbytes) + 3) & ~3L, SEEK_CUR) == 0;
}

}  // namespace
}  // namespace cir

using namespace cir;

struct cir_xtc {
    FILE* f = nullptr;
    int natoms = 0;
    int frame = 0;                // index of the next frame
    std::vector<int64_t> offset;  // byte offset of every frame, built on the first seek
    ~cir_xtc() {
        if (f) std::fclose(f);
    }
};

namespace {
// Byte offsets of the frames of an open file, from the frame headers and the
// byte counts of the coordinate blocks (nothing is decoded); leaves the file
// position undefined.
int frame_offsets(FILE* f, std::vector<int64_t>& off) {
    off.clear();
    if (std::fseek(f, 0, SEEK_SET) != 0) return fail(CIR_ERR_IO, "xtc: seek failed");
    XdrIn x{f};
    for (;;) {
        const int64_t at = int64_t(std::ftell(f));
        int nat, step;
        float t, bx[9];
        if (!read_header(x, &nat, &step, &t, bx)) break;
        if (!skip_coords(x, nat)) return fail(CIR_ERR_IO, "xtc: corrupt frame " + std::to_string(off.size()));
        off.push_back(at);
    }
    std::clearerr(f);
    return CIR_OK;
}
}  // namespace

extern "C" int cir_xtc_offsets(const char* path, int nframes, int64_t* offsets) try {
    if (!path || nframes < 0 || (nframes > 0 && !offsets)) return fail(CIR_ERR_ARG, "cir_xtc_offsets: bad argument");
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(CIR_ERR_IO, std::string("could not open ") + path);
    std::vector<int64_t> off;
    const int st = frame_offsets(f, off);
    std::fclose(f);
    if (st != CIR_OK) return st;
    if (int64_t(off.size()) < nframes) return fail(CIR_ERR_FRAME, "xtc: frame " + std::to_string(off.size()) + " missing");
    for (int k = 0; k < nframes; ++k) offsets[k] = off[k];
    return CIR_OK;
} catch (...) {
    return cir::host_exception("cir_xtc_offsets");
}

extern "C" int cir_xtc_seek(cir_xtc* r, int frame) try {
    if (!r || frame < 0) return fail(CIR_ERR_ARG, "cir_xtc_seek: bad argument");
    if (r->offset.empty()) CIR_TRY(frame_offsets(r->f, r->offset));
    if (frame >= int(r->offset.size()))
        return fail(CIR_ERR_FRAME, "xtc: frame " + std::to_string(frame) + " missing (" +
                                       std::to_string(r->offset.size()) + " frames)");
    if (std::fseek(r->f, long(r->offset[frame]), SEEK_SET) != 0) return fail(CIR_ERR_IO, "xtc: seek failed");
    r->frame = frame;
    return CIR_OK;
} catch (...) {
    return cir::host_exception("cir_xtc_seek");
}

extern "C" int cir_xtc_open(const char* path, int natoms, cir_xtc** out) try {
    if (!path || !out || natoms < 0) return fail(CIR_ERR_ARG, "cir_xtc_open: bad argument");
    *out = nullptr;
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(CIR_ERR_IO, std::string("could not open ") + path);
    cir_xtc* r = new cir_xtc();
    r->f = f;
    r->natoms = natoms;
    *out = r;
    return CIR_OK;
} catch (...) {
    return cir::host_exception("cir_xtc_open");
}

extern "C" int cir_xtc_next(cir_xtc* r, int count, float* xyz, float* box9, float* times) try {
    if (!r || count < 0 || (count > 0 && !xyz)) return fail(CIR_ERR_ARG, "cir_xtc_next: bad argument");
    XdrIn x{r->f};
    for (int k = 0; k < count; ++k) {
        int nat, step;
        float t, bx[9];
        if (!read_header(x, &nat, &step, &t, bx)) return fail(CIR_ERR_FRAME, "xtc: frame " + std::to_string(r->frame) + " missing");
        if (nat != r->natoms) return fail(CIR_ERR_IO, "xtc: atom count differs from the expected " + std::to_string(r->natoms));
        if (!read_coords(x, nat, xyz + size_t(k) * 3 * nat))
            return fail(CIR_ERR_IO, "xtc: corrupt coordinates in frame " + std::to_string(r->frame));
        if (box9)
            for (int q = 0; q < 9; ++q) box9[9 * size_t(k) + q] = bx[q];
        if (times) times[k] = t;
        ++r->frame;
    }
    return CIR_OK;
} catch (...) {
    return cir::host_exception("cir_xtc_next");
}

extern "C" void cir_xtc_close(cir_xtc* r) { delete r; }

extern "C" int cir_xtc_times(const char* path, int nframes, float* times, float* box9) try {
    if (!path || nframes < 0) return fail(CIR_ERR_ARG, "cir_xtc_times: bad argument");
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(CIR_ERR_IO, std::string("could not open ") + path);
    XdrIn x{f};
    int na = -1;
    for (int k = 0; k < nframes; ++k) {
        int nat, step;
        float t, bx[9];
        if (!read_header(x, &nat, &step, &t, bx)) {
            std::fclose(f);
            return fail(CIR_ERR_FRAME, "xtc: frame " + std::to_string(k) + " missing");
        }
        if ((na >= 0 && nat != na) || !skip_coords(x, nat)) {
            std::fclose(f);
            return fail(CIR_ERR_IO, "xtc: corrupt frame " + std::to_string(k));
        }
        na = nat;
        if (times) times[k] = t;
        if (box9)
            for (int q = 0; q < 9; ++q) box9[9 * size_t(k) + q] = bx[q];
    }
    std::fclose(f);
    return CIR_OK;
} catch (...) {
    return cir::host_exception("cir_xtc_times");
}

extern "C" int cir_xtc_precision(const char* path, int frame, float* precision) try {
    if (!path || frame < 0 || !precision) return fail(CIR_ERR_ARG, "cir_xtc_precision: bad argument");
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(CIR_ERR_IO, std::string("could not open ") + path);
    XdrIn x{f};
    int st = CIR_OK;
    for (int k = 0;; ++k) {
        int nat, step;
        float t, bx[9];
        if (!read_header(x, &nat, &step, &t, bx)) {
            st = fail(CIR_ERR_FRAME, "xtc: frame " + std::to_string(k) + " missing");
            break;
        }
        if (k < frame) {
            if (!skip_coords(x, nat)) {
                st = fail(CIR_ERR_IO, "xtc: corrupt frame " + std::to_string(k));
                break;
            }
            continue;
        }
        const int lsize = x.i32();
        if (!x.ok || lsize != nat) {
            st = fail(CIR_ERR_IO, "xtc: corrupt frame " + std::to_string(k));
            break;
        }
        if (nat <= 9) {
            *precision = -1.0f;  // stored uncompressed, as xdrfile reports it
        } else {
            const float pr = x.f32();
            if (!x.ok || !(pr > 0.0f)) {
                st = fail(CIR_ERR_IO, "xtc: corrupt precision in frame " + std::to_string(k));
                break;
            }
            *precision = pr;
        }
        break;
    }
    std::fclose(f);
    return st;
} catch (...) {
    return cir::host_exception("cir_xtc_precision");
}

extern "C" int cir_xtc_scan(const char* path, int* natoms, int* nframes) try {
    if (!path || !natoms || !nframes) return fail(CIR_ERR_ARG, "cir_xtc_scan: null argument");
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(CIR_ERR_IO, std::string("could not open ") + path);
    XdrIn x{f};
    // every atom takes at least two bits of the file: a larger count in a
    // header is corrupt (and would size the decode buffer from garbage)
    long long fsize = 0;
    if (std::fseek(f, 0, SEEK_END) == 0) fsize = std::ftell(f);
    std::rewind(f);
    int n = 0, na = -1;
    std::vector<float> buf;
    for (;;) {
        int nat, step;
        float t, box[9];
        if (!read_header(x, &nat, &step, &t, box)) break;
        if ((na >= 0 && nat != na) || nat > 4 * fsize + 9) {
            std::fclose(f);
            return fail(CIR_ERR_IO, "xtc: atom count changes between frames or exceeds the file");
        }
        na = nat;
        buf.resize(3 * size_t(nat));
        if (!read_coords(x, nat, buf.data())) {
            std::fclose(f);
            return fail(CIR_ERR_IO, "xtc: corrupt coordinates in frame " + std::to_string(n));
        }
        ++n;
    }
    const bool clean = std::feof(f) != 0;
    std::fclose(f);
    if (!clean && n == 0) return fail(CIR_ERR_IO, std::string(path) + " is not an XTC file");
    *natoms = (na < 0) ? 0 : na;
    *nframes = n;
    return CIR_OK;
} catch (...) {
    return cir::host_exception("cir_xtc_scan");
}

namespace {

// Frames first..first+count-1; box gets bs floats per frame: 3 (the
// diagonal) or 9 (the full matrix, rows a, b, c).
int read_frames(const char* path, int first, int count, int natoms, float* xyz, float* box, int bs, float* times,
                int* steps) {
    if (!path || first < 0 || count < 0 || !xyz) return fail(CIR_ERR_ARG, "cir_xtc_read: bad argument");
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(CIR_ERR_IO, std::string("could not open ") + path);
    XdrIn x{f};
    int got = 0;
    for (int fr = 0; got < count; ++fr) {
        int nat, step;
        float t, bx[9];
        if (!read_header(x, &nat, &step, &t, bx)) {
            std::fclose(f);
            return fail(CIR_ERR_FRAME, "xtc: frame " + std::to_string(fr) + " missing");
        }
        if (nat != natoms) {
            std::fclose(f);
            return fail(CIR_ERR_IO, "xtc: atom count differs from the expected " + std::to_string(natoms));
        }
        // frames before the range are skipped by their byte counts
        const bool ok = (fr >= first) ? read_coords(x, nat, xyz + size_t(got) * 3 * natoms) : skip_coords(x, nat);
        if (!ok) {
            std::fclose(f);
            return fail(CIR_ERR_IO, "xtc: corrupt coordinates in frame " + std::to_string(fr));
        }
        if (fr >= first) {
            if (box && bs == 3)  // the diagonal, as the reference uses it (calcIR.cu:404-405)
                for (int k = 0; k < 3; ++k) box[3 * size_t(got) + k] = bx[4 * k];
            if (box && bs == 9)
                for (int k = 0; k < 9; ++k) box[9 * size_t(got) + k] = bx[k];
            if (times) times[got] = t;
            if (steps) steps[got] = step;
            ++got;
        }
    }
    std::fclose(f);
    return CIR_OK;
}

int write_frames(const char* path, int natoms, int nframes, const float* xyz, const float* box, int bs,
                 const float* times, float precision, int append) {
    if (!path || natoms < 0 || nframes < 0 || (nframes > 0 && !xyz) || !(precision > 0.0f))
        return fail(CIR_ERR_ARG, "cir_xtc_write: bad argument");
    FILE* f = std::fopen(path, append ? "ab" : "wb");
    if (!f) return fail(CIR_ERR_IO, std::string("could not open ") + path + " for writing");
    for (int fr = 0; fr < nframes; ++fr) {
        XdrOut x;
        x.i32(kXtcMagic);
        x.i32(natoms);
        x.i32(fr);
        x.f32(times ? times[fr] : float(fr));
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                float v = 0.0f;
                if (box && bs == 3 && r == c) v = box[3 * size_t(fr) + r];
                if (box && bs == 9) v = box[9 * size_t(fr) + 3 * r + c];
                x.f32(v);
            }
        if (!write_coords(x, natoms, xyz + size_t(fr) * 3 * natoms, precision)) {
            std::fclose(f);
            return fail(CIR_ERR_ARG, "xtc: coordinates out of range at this precision");
        }
        if (std::fwrite(x.b.data(), 1, x.b.size(), f) != x.b.size()) {
            std::fclose(f);
            return fail(CIR_ERR_IO, std::string("write to ") + path + " failed");
        }
    }
    std::fclose(f);
    return CIR_OK;
}

}  // namespace

extern "C" int cir_xtc_read(const char* path, int first, int count, int natoms, float* xyz, float* box,
                            float* times, int* steps) try {
    return read_frames(path, first, count, natoms, xyz, box, 3, times, steps);
} catch (...) {
    return cir::host_exception("cir_xtc_read");
}

extern "C" int cir_xtc_read_full(const char* path, int first, int count, int natoms, float* xyz, float* box9,
                                 float* times, int* steps) try {
    return read_frames(path, first, count, natoms, xyz, box9, 9, times, steps);
} catch (...) {
    return cir::host_exception("cir_xtc_read_full");
}

extern "C" int cir_xtc_write(const char* path, int natoms, int nframes, const float* xyz, const float* box,
                             const float* times, float precision, int append) try {
    return write_frames(path, natoms, nframes, xyz, box, 3, times, precision, append);
} catch (...) {
    return cir::host_exception("cir_xtc_write");
}

extern "C" int cir_xtc_write_full(const char* path, int natoms, int nframes, const float* xyz, const float* box9,
                                  const float* times, float precision, int append) try {
    return write_frames(path, natoms, nframes, xyz, box9, 9, times, precision, append);
} catch (...) {
    return cir::host_exception("cir_xtc_write_full");
}

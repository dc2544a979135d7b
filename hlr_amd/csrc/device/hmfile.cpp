// hmfile.cpp -- on-disk format of a flattened H- or TLR-matrix.
//
// Takes the place of hlr::io::hpro::write / read (utils/io.hh:59-80): the
// reference stores a TMatrix through HLIBpro's own file formats, which this
// library cannot read or write.  The file holds what hmvm_plan_create takes
// -- the leaf records and the FP64 factor arena -- so a stored matrix comes
// back through the same validated plan build, in any storage precision,
// without running the generator again.
//
// Layout (little-endian, 8-byte fields):
//   [0, 8)    magic "HMVMHMAT"
//   [8, 12)   version (1)            [12, 16)  zero
//   [16, 48)  nrows, ncols, nblocks, arena_len (int64)
//   [48, 56)  checksum of the block records
//   [56, 64)  checksum of the arena
//   [64, 72)  checksum of bytes [0, 64)
//   [72, ...) nblocks records of 9 int64: row_ofs, col_ofs, nrows, ncols,
//             rank, kind, off_a, off_b, off_s
//   then      arena_len FP64 values
// A checksum folds the section's 8-byte words in order:
// h = ((h ^ w) * 0x9E3779B97F4A7C15) ^ (h >> 29), h0 = 0x484D564D.

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "hmvm.h"

namespace {

static_assert(__BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__, "records and arena are written in host byte order");

constexpr char kMagic[8] = {'H', 'M', 'V', 'M', 'H', 'M', 'A', 'T'};
constexpr uint32_t kVersion = 1;
constexpr int kHeaderBytes = 72;
constexpr int kRecordWords = 9;
constexpr int64_t kChunkWords = int64_t(1) << 20;   // 8 MB per read / write

struct Sum {
    uint64_t h = 0x484D564Dull;
    void add(const uint64_t* w, int64_t n) {
        uint64_t x = h;
        for (int64_t i = 0; i < n; ++i) x = ((x ^ w[i]) * 0x9E3779B97F4A7C15ull) ^ (x >> 29);
        h = x;
    }
};

struct File {
    std::FILE* f = nullptr;
    File(const char* path, const char* mode) : f(std::fopen(path, mode)) {}
    ~File() { if (f) std::fclose(f); }
};

void put64(unsigned char* p, uint64_t v) {
    for (int i = 0; i < 8; ++i) p[i] = (unsigned char)(v >> (8 * i));
}
uint64_t get64(const unsigned char* p) {
    uint64_t v = 0;
    for (int i = 0; i < 8; ++i) v |= uint64_t(p[i]) << (8 * i);
    return v;
}

void block_words(const hmvm_block& b, uint64_t w[kRecordWords]) {
    const int64_t v[kRecordWords] = {b.row_ofs, b.col_ofs, b.nrows, b.ncols, b.rank,
                                     b.kind,    b.off_a,   b.off_b, b.off_s};
    for (int i = 0; i < kRecordWords; ++i) w[i] = uint64_t(v[i]);
}

bool words_block(const uint64_t w[kRecordWords], hmvm_block& b) {
    auto i32 = [](uint64_t x, int32_t& out) {
        const int64_t v = int64_t(x);
        if (v < INT32_MIN || v > INT32_MAX) return false;
        out = int32_t(v);
        return true;
    };
    b.row_ofs = int64_t(w[0]);
    b.col_ofs = int64_t(w[1]);
    b.off_a = int64_t(w[6]);
    b.off_b = int64_t(w[7]);
    b.off_s = int64_t(w[8]);
    return i32(w[2], b.nrows) && i32(w[3], b.ncols) && i32(w[4], b.rank) && i32(w[5], b.kind);
}

struct Header {
    int64_t nrows = 0, ncols = 0, nblocks = 0, arena_len = 0;
    uint64_t sum_blocks = 0, sum_arena = 0;
};

// read and check the header
int read_header(std::FILE* f, Header& h) {
    unsigned char buf[kHeaderBytes];
    if (std::fread(buf, 1, kHeaderBytes, f) != size_t(kHeaderBytes)) return HMVM_ERR_IO;
    if (std::memcmp(buf, kMagic, 8) != 0) return HMVM_ERR_IO;
    uint64_t words[8];
    for (int i = 0; i < 8; ++i) words[i] = get64(buf + 8 * i);
    Sum s;
    s.add(words, 8);
    if (s.h != get64(buf + 64)) return HMVM_ERR_CORRUPT;
    const uint32_t version = uint32_t(buf[8]) | uint32_t(buf[9]) << 8 | uint32_t(buf[10]) << 16 |
                             uint32_t(buf[11]) << 24;
    if (version != kVersion) return HMVM_ERR_IO;
    h.nrows = int64_t(get64(buf + 16));
    h.ncols = int64_t(get64(buf + 24));
    h.nblocks = int64_t(get64(buf + 32));
    h.arena_len = int64_t(get64(buf + 40));
    h.sum_blocks = get64(buf + 48);
    h.sum_arena = get64(buf + 56);
    if (h.nrows < 0 || h.ncols < 0 || h.nblocks < 0 || h.arena_len < 0) return HMVM_ERR_CORRUPT;
    return HMVM_OK;
}

} // namespace

extern "C" int hmvm_hmatrix_write(const char* path, const hmvm_block* blocks, int64_t nblocks, int64_t nrows,
                                  int64_t ncols, const double* arena, int64_t arena_len) {
    if (!path || nblocks < 0 || nrows < 0 || ncols < 0 || arena_len < 0 || (nblocks > 0 && !blocks) ||
        (arena_len > 0 && !arena))
        return HMVM_ERR_INVALID_ARG;
    static_assert(sizeof(double) == 8, "FP64 arena");
    // block records and both section checksums
    std::vector<uint64_t> rec(size_t(nblocks) * kRecordWords);
    for (int64_t b = 0; b < nblocks; ++b) block_words(blocks[b], rec.data() + b * kRecordWords);
    Sum sb, sa;
    sb.add(rec.data(), int64_t(rec.size()));
    std::vector<uint64_t> words(size_t(std::min<int64_t>(arena_len, kChunkWords)));
    for (int64_t i = 0; i < arena_len; i += kChunkWords) {
        const int64_t n = std::min<int64_t>(kChunkWords, arena_len - i);
        std::memcpy(words.data(), arena + i, size_t(n) * 8);
        sa.add(words.data(), n);
    }
    unsigned char hdr[kHeaderBytes] = {};
    std::memcpy(hdr, kMagic, 8);
    for (int i = 0; i < 4; ++i) hdr[8 + i] = (unsigned char)(kVersion >> (8 * i));
    put64(hdr + 16, uint64_t(nrows));
    put64(hdr + 24, uint64_t(ncols));
    put64(hdr + 32, uint64_t(nblocks));
    put64(hdr + 40, uint64_t(arena_len));
    put64(hdr + 48, sb.h);
    put64(hdr + 56, sa.h);
    uint64_t hw[8];
    for (int i = 0; i < 8; ++i) hw[i] = get64(hdr + 8 * i);
    Sum sh;
    sh.add(hw, 8);
    put64(hdr + 64, sh.h);
    bool ok;
    {
        File out(path, "wb");
        if (!out.f) return HMVM_ERR_IO;
        ok = std::fwrite(hdr, 1, kHeaderBytes, out.f) == size_t(kHeaderBytes);
        // records and arena as little-endian words (the host byte order)
        if (ok && !rec.empty()) ok = std::fwrite(rec.data(), 8, rec.size(), out.f) == rec.size();
        for (int64_t i = 0; ok && i < arena_len; i += kChunkWords) {
            const int64_t n = std::min<int64_t>(kChunkWords, arena_len - i);
            ok = std::fwrite(arena + i, 8, size_t(n), out.f) == size_t(n);
        }
        ok = std::fclose(out.f) == 0 && ok;
        out.f = nullptr;
    }
    if (!ok) {
        std::remove(path);
        return HMVM_ERR_IO;
    }
    return HMVM_OK;
}

extern "C" int hmvm_hmatrix_read_info(const char* path, int64_t* nrows, int64_t* ncols, int64_t* nblocks,
                                      int64_t* arena_len) {
    if (!path) return HMVM_ERR_INVALID_ARG;
    File in(path, "rb");
    if (!in.f) return HMVM_ERR_IO;
    Header h;
    const int rc = read_header(in.f, h);
    if (rc != HMVM_OK) return rc;
    if (nrows) *nrows = h.nrows;
    if (ncols) *ncols = h.ncols;
    if (nblocks) *nblocks = h.nblocks;
    if (arena_len) *arena_len = h.arena_len;
    return HMVM_OK;
}

extern "C" int hmvm_hmatrix_read(const char* path, hmvm_block* blocks, int64_t nblocks, double* arena,
                                 int64_t arena_len) {
    if (!path || nblocks < 0 || arena_len < 0 || (nblocks > 0 && !blocks) || (arena_len > 0 && !arena))
        return HMVM_ERR_INVALID_ARG;
    File in(path, "rb");
    if (!in.f) return HMVM_ERR_IO;
    Header h;
    int rc = read_header(in.f, h);
    if (rc != HMVM_OK) return rc;
    if (h.nblocks != nblocks || h.arena_len != arena_len) return HMVM_ERR_INVALID_ARG;
    std::vector<uint64_t> rec(size_t(nblocks) * kRecordWords);
    if (!rec.empty() && std::fread(rec.data(), 8, rec.size(), in.f) != rec.size()) return HMVM_ERR_IO;
    Sum sb, sa;
    sb.add(rec.data(), int64_t(rec.size()));
    if (sb.h != h.sum_blocks) return HMVM_ERR_CORRUPT;
    for (int64_t b = 0; b < nblocks; ++b)
        if (!words_block(rec.data() + b * kRecordWords, blocks[b])) return HMVM_ERR_CORRUPT;
    std::vector<uint64_t> words(size_t(std::min<int64_t>(arena_len, kChunkWords)));
    for (int64_t i = 0; i < arena_len; i += kChunkWords) {
        const int64_t n = std::min<int64_t>(kChunkWords, arena_len - i);
        if (std::fread(words.data(), 8, size_t(n), in.f) != size_t(n)) return HMVM_ERR_IO;
        sa.add(words.data(), n);
        std::memcpy(arena + i, words.data(), size_t(n) * 8);
    }
    if (sa.h != h.sum_arena) return HMVM_ERR_CORRUPT;
    // nothing may follow the arena
    if (std::fgetc(in.f) != EOF) return HMVM_ERR_CORRUPT;
    return HMVM_OK;
}

extern "C" int hmvm_plan_create_from_file(const char* path, const hmvm_plan_opts* opts, int device,
                                          hmvm_plan** out) {
    if (!path || !out) return HMVM_ERR_INVALID_ARG;
    *out = nullptr;
    int64_t nrows, ncols, nblocks, arena_len;
    int rc = hmvm_hmatrix_read_info(path, &nrows, &ncols, &nblocks, &arena_len);
    if (rc != HMVM_OK) return rc;
    std::vector<hmvm_block> blocks;
    std::vector<double> arena;
    try {
        blocks.resize(size_t(nblocks));
        arena.resize(size_t(arena_len));
    } catch (...) {
        return HMVM_ERR_ALLOC;
    }
    rc = hmvm_hmatrix_read(path, blocks.data(), nblocks, arena.data(), arena_len);
    if (rc != HMVM_OK) return rc;
    return hmvm_plan_create(blocks.data(), nblocks, nrows, ncols, arena.data(), arena_len, opts, device, out);
}

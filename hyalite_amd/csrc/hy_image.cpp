// On-disk database image: the scheme, the builder fields and the residues in database order, in
// one little-endian file with a checksum. Loading re-runs the build on those inputs, whose
// per-residue work (the lane-interleaved layouts) happens on the device, so a saved database
// comes back with one sequential read, one upload of its residues and a few gather kernels,
// instead of re-reading and encoding its source sequences. The reference keeps sequences in
// memory only (src/database.rs:674-677); this format is this backend's own.
//
// Layout (all integers little-endian):
//   magic "HYDBIMG1" | u32 version (1) | u32 alphabet_len | i32 gap_open | i32 gap_ext
//   | u32 mode | u32 search_type | u64 traceback_max_bytes | u64 max_query_len | u64 n
//   | u64 total_residues | i32 matrix[alphabet_len^2] | u64 offsets[n + 1] | u8 residues[total]
//   | u64 checksum of every preceding byte: h = 0xcbf29ce484222325, then for each little-endian
//     u64 word w of those bytes (the last one zero-padded) h = (h ^ w) * 0x100000001b3 mod 2^64
//     (the FNV-1a step applied to words, so loading hashes 8 bytes per multiply)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <thread>
#include <memory>
#include <vector>

#include "../../include/hyalite_hip.h"
#include "hy_internal.h"
#include "hy_state.h"

namespace {

constexpr char kMagic[8] = {'H', 'Y', 'D', 'B', 'I', 'M', 'G', '1'};
constexpr uint32_t kVersion = 1;

// The word-wise checksum over a byte stream fed in pieces of any length.
struct Fnv {
    uint64_t h = 0xcbf29ce484222325ull;
    uint64_t pend = 0;   // bytes of the word being assembled
    unsigned npend = 0;
    void word(uint64_t w) { h = (h ^ w) * 0x100000001b3ull; }
    void add(const void* p, size_t n) {
        const uint8_t* b = static_cast<const uint8_t*>(p);
        while (n && npend) {
            pend |= (uint64_t)*b++ << (8 * npend++);
            --n;
            if (npend == 8) { word(pend); pend = 0; npend = 0; }
        }
        for (; n >= 8; n -= 8, b += 8) {
            uint64_t w;
            std::memcpy(&w, b, 8);
            word(w);
        }
        for (; n; --n) pend |= (uint64_t)*b++ << (8 * npend++);
    }
    uint64_t digest() const { return npend ? (h ^ pend) * 0x100000001b3ull : h; }
};

struct Header {
    char magic[8];
    uint32_t version, al;
    int32_t go, ge;
    uint32_t mode, search_type;
    uint64_t budget, mql, n, total;
};
static_assert(sizeof(Header) == 64, "packed image header");

struct File {
    std::FILE* f = nullptr;
    ~File() { if (f) std::fclose(f); }
};

}  // namespace

extern "C" int hy_db_save(const hy_db* db, const char* path) {
    if (!db || !path) return HY_E_NULL_ARGUMENT;
    if (hipSetDevice(db->device) != hipSuccess) return HY_E_DEVICE;
    std::vector<uint8_t> res(db->total_residues);
    if (!res.empty() && hipMemcpy(res.data(), db->d_flat, res.size(), hipMemcpyDeviceToHost) !=
                            hipSuccess)
        return HY_E_DEVICE;
    Header h{};
    std::memcpy(h.magic, kMagic, 8);
    h.version = kVersion;
    h.al = db->al;
    h.go = db->go;
    h.ge = db->ge;
    h.mode = db->mode;
    h.search_type = db->search_type;
    h.budget = db->traceback_max_bytes;
    h.mql = db->mql;
    h.n = db->n;
    h.total = db->total_residues;
    File out;
    out.f = std::fopen(path, "wb");
    if (!out.f) return HY_E_INVALID_ARGUMENT;
    Fnv sum;
    auto put = [&](const void* p, size_t n) {
        sum.add(p, n);
        return n == 0 || std::fwrite(p, 1, n, out.f) == n;
    };
    bool ok = put(&h, sizeof(h)) &&
              put(db->h_matrix.data(), db->h_matrix.size() * sizeof(int32_t)) &&
              put(db->h_off.data(), db->h_off.size() * sizeof(uint64_t)) &&
              put(res.data(), res.size());
    const uint64_t digest = sum.digest();
    ok = ok && std::fwrite(&digest, 1, sizeof(digest), out.f) == sizeof(digest);
    ok = std::fclose(out.f) == 0 && ok;
    out.f = nullptr;
    return ok ? HY_OK : HY_E_INVALID_ARGUMENT;
}

extern "C" int hy_db_load(const char* path, uint32_t backend, hy_db** out, int64_t* err_detail) {
    if (!path || !out) return HY_E_NULL_ARGUMENT;
    std::lock_guard<std::recursive_mutex> api(hy::api_lock());
    *out = nullptr;
    if (err_detail) *err_detail = 0;
    File in;
    in.f = std::fopen(path, "rb");
    if (!in.f) return HY_E_INVALID_ARGUMENT;
    if (std::fseek(in.f, 0, SEEK_END) != 0) return HY_E_INVALID_ARGUMENT;
    const long size = std::ftell(in.f);
    if (size < (long)sizeof(Header) + 8 || std::fseek(in.f, 0, SEEK_SET) != 0)
        return HY_E_BAD_IMAGE;
    Header h{};
    if (std::fread(&h, 1, sizeof(h), in.f) != sizeof(h)) return HY_E_BAD_IMAGE;
    if (std::memcmp(h.magic, kMagic, 8) != 0 || h.version != kVersion) return HY_E_BAD_IMAGE;
    // Every section's size follows from the header; the file must hold exactly them.
    // Any alphabet the build accepts can be saved (Scoring::new takes alphabets above 256, whose
    // larger symbols simply never occur in a byte sequence), so the load takes it too; the matrix
    // section is bounded by the file size before the exact-size check.
    const uint64_t mat_n = (uint64_t)h.al * h.al;
    if (h.al == 0 || mat_n > (uint64_t)size / 4 || h.n == 0 || h.n > 0xFFFFFFF0ull)
        return HY_E_BAD_IMAGE;
    const uint64_t expect = sizeof(Header) + mat_n * 4 + (h.n + 1) * 8 + h.total + 8;
    if (h.total > (uint64_t)size || (uint64_t)size != expect) return HY_E_BAD_IMAGE;
    std::vector<int32_t> mat(mat_n);
    std::vector<uint64_t> off(h.n + 1);
    // Every residue byte is read from the file, so the buffer is left uninitialised.
    std::unique_ptr<uint8_t[]> res(new (std::nothrow) uint8_t[std::max<uint64_t>(h.total, 1)]);
    if (!res) return HY_E_OUT_OF_MEMORY;
    uint64_t digest = 0;
    if (std::fread(mat.data(), 4, mat_n, in.f) != mat_n ||
        std::fread(off.data(), 8, h.n + 1, in.f) != h.n + 1 ||
        (h.total && std::fread(res.get(), 1, h.total, in.f) != h.total) ||
        std::fread(&digest, 8, 1, in.f) != 1)
        return HY_E_BAD_IMAGE;
    // Offsets: from 0, non-decreasing, ending at the residue count (the build below indexes the
    // residues with them).
    if (off[0] != 0 || off[h.n] != h.total) return HY_E_BAD_IMAGE;
    std::vector<const uint8_t*> ptrs(h.n);
    std::vector<size_t> lens(h.n);
    for (uint64_t k = 0; k < h.n; ++k) {
        if (off[k + 1] < off[k]) return HY_E_BAD_IMAGE;
        ptrs[k] = res.get() + off[k];
        lens[k] = (size_t)(off[k + 1] - off[k]);
    }
    // The checksum (one serial multiply chain over the file) runs beside the build, which only
    // reads the same buffers; a mismatch discards the build's result, whatever it was.
    bool sum_ok = false;
    auto checksum = [&]() {
        Fnv sum;
        sum.add(&h, sizeof(h));
        sum.add(mat.data(), mat_n * 4);
        sum.add(off.data(), (h.n + 1) * 8);
        sum.add(res.get(), h.total);
        sum_ok = sum.digest() == digest;
    };
    std::thread check;
    try {
        check = std::thread(checksum);
    } catch (...) {  // no second thread: check it here, before the build
        checksum();
        if (!sum_ok) return HY_E_BAD_IMAGE;
    }
    // The build validates the scheme, every symbol, the width proof and the traceback budget
    // again, exactly as for sequences given in memory.
    int64_t detail = 0;
    const int rc = hy_db_build(ptrs.data(), lens.data(), h.n, HY_HAVE_ALL, mat.data(), mat_n,
                               h.al, h.go, h.ge, h.mode, h.search_type, h.budget, h.mql, backend,
                               out, &detail);
    if (check.joinable()) check.join();
    if (!sum_ok) {
        if (*out) hy_db_free(*out);
        *out = nullptr;
        return HY_E_BAD_IMAGE;
    }
    if (err_detail) *err_detail = detail;
    return rc;
}

extern "C" int hy_db_scoring(const hy_db* db, int32_t* matrix, uint32_t* alphabet_len,
                             int32_t* gap_open, int32_t* gap_ext, uint64_t* traceback_max_bytes) {
    if (!db) return HY_E_NULL_ARGUMENT;
    if (matrix) std::memcpy(matrix, db->h_matrix.data(), db->h_matrix.size() * sizeof(int32_t));
    if (alphabet_len) *alphabet_len = db->al;
    if (gap_open) *gap_open = db->go;
    if (gap_ext) *gap_ext = db->ge;
    if (traceback_max_bytes) *traceback_max_bytes = db->traceback_max_bytes;
    return HY_OK;
}

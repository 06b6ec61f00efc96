// Checksummed binary files of the library (world snapshots, warm-start network files):
// a fixed header (8-byte magic, version), little-endian sections in a fixed order, and
// a trailing 64-bit FNV-1a hash of every byte before it.  A reader rejects a file whose
// magic, version, sizes or hash do not match.
#pragma once
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

namespace axm {

inline uint64_t fnv1a(uint64_t h, const void* p, size_t len) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < len; ++i) {
        h ^= b[i];
        h *= 0x100000001B3ull;
    }
    return h;
}
constexpr uint64_t FNV_SEED = 0xCBF29CE484222325ull;

class BinWriter {
public:
    explicit BinWriter(const char* path) : f_(std::fopen(path, "wb")) {}
    ~BinWriter() {
        if (f_) std::fclose(f_);
    }
    bool ok() const { return f_ != nullptr && ok_; }
    void put(const void* p, size_t len) {
        if (!ok() || len == 0) return;
        h_ = fnv1a(h_, p, len);
        ok_ = std::fwrite(p, 1, len, f_) == len;
    }
    template <class T> void pod(const T& v) { put(&v, sizeof(T)); }
    template <class T> void vec(const std::vector<T>& v) { put(v.data(), sizeof(T) * v.size()); }
    // writes the hash and closes; false if any write failed
    bool finish() {
        if (!ok()) return false;
        const uint64_t h = h_;
        ok_ = std::fwrite(&h, 1, 8, f_) == 8;
        ok_ = (std::fclose(f_) == 0) && ok_;
        f_ = nullptr;
        return ok_;
    }

private:
    std::FILE* f_;
    uint64_t h_ = FNV_SEED;
    bool ok_ = true;
};

class BinReader {
public:
    explicit BinReader(const char* path) : f_(std::fopen(path, "rb")) {}
    ~BinReader() {
        if (f_) std::fclose(f_);
    }
    bool ok() const { return f_ != nullptr && ok_; }
    void get(void* p, size_t len) {
        if (!ok() || len == 0) return;
        ok_ = std::fread(p, 1, len, f_) == len;
        if (ok_) h_ = fnv1a(h_, p, len);
    }
    template <class T> void pod(T& v) { get(&v, sizeof(T)); }
    template <class T> void vec(std::vector<T>& v, size_t n) {
        v.resize(n);
        get(v.data(), sizeof(T) * n);
    }
    // reads the trailing hash; true when it matches and nothing follows it
    bool finish() {
        if (!ok()) return false;
        uint64_t h = 0;
        const uint64_t want = h_;
        if (std::fread(&h, 1, 8, f_) != 8 || h != want) return false;
        unsigned char extra;
        return std::fread(&extra, 1, 1, f_) == 0;
    }

private:
    std::FILE* f_;
    uint64_t h_ = FNV_SEED;
    bool ok_ = true;
};

}  // namespace axm

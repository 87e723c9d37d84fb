/*
 * fileio.cpp - The formats of fileio.h with 64-bit file offsets (POSIX).
 */
#include "fileio.h"

#include <cstdio>
#include <sys/stat.h>

namespace {
    int64_t sizeOf(FILE* f) {
        struct stat st;
        return fstat(fileno(f), &st) == 0 ? static_cast<int64_t>(st.st_size) : -1;
    }

    struct File {
        FILE* f;
        explicit File(FILE* file) : f(file) {}
        ~File() {
            if (f) {
                fclose(f);
            }
        }
        bool close() {
            const bool ok = fclose(f) == 0;
            f = nullptr;
            return ok;
        }
    };
}

bool readWholeFile(const char* path, std::vector<uint8_t>& out, size_t padding) {
    File file(fopen(path, "rb"));
    if (!file.f) {
        return false;
    }
    const int64_t size = sizeOf(file.f);
    if (size < 0) {
        return false;
    }
    out.assign(static_cast<size_t>(size) + padding, 0);
    return fread(out.data(), 1, static_cast<size_t>(size), file.f) == static_cast<size_t>(size);
}

bool readSuffixArray(const char* path, std::vector<int32_t>& out) {
    File file(fopen(path, "rb"));
    if (!file.f) {
        return false;
    }
    const int64_t size = sizeOf(file.f);
    if (size < 0 || size % 4 != 0) {
        return false;
    }
    out.resize(static_cast<size_t>(size / 4));
    return fread(out.data(), 4, out.size(), file.f) == out.size();
}

bool writeWholeFile(const char* path, const void* data, size_t size) {
    File file(fopen(path, "wb"));
    if (!file.f) {
        return false;
    }
    const bool written = fwrite(data, 1, size, file.f) == size;
    return file.close() && written;
}

bool readBwtFile(const char* path, std::vector<uint8_t>& out, int32_t& primary) {
    File file(fopen(path, "rb"));
    if (!file.f) {
        return false;
    }
    const int64_t size = sizeOf(file.f);
    int32_t n = -1;
    primary = 0;
    if (size < 4 || fread(&n, 4, 1, file.f) != 1 || n < 0) {
        return false;
    }
    if (n == 0) {
        out.clear();
        return size == 4;
    }
    if (size != 8 + static_cast<int64_t>(n) || fread(&primary, 4, 1, file.f) != 1) {
        return false;
    }
    out.resize(static_cast<size_t>(n));
    return fread(out.data(), 1, out.size(), file.f) == out.size();
}

bool writeBwtFile(const char* path, const uint8_t* bwt, int32_t n, int32_t primary) {
    File file(fopen(path, "wb"));
    if (!file.f) {
        return false;
    }
    bool ok = fwrite(&n, 4, 1, file.f) == 1;
    if (ok && n > 0) {
        ok = fwrite(&primary, 4, 1, file.f) == 1 && fwrite(bwt, 1, static_cast<size_t>(n), file.f) == static_cast<size_t>(n);
    }
    return file.close() && ok;
}

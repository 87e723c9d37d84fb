/*
 * fileio.h - divsufsort's file formats (the reference's src/common/fileio.h:16-29):
 * a suffix array file is n little-endian int32; a BWT file is one block: int32 n, int32 primary
 * (omitted when n = 0), then the n bytes.
 */
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

// The whole file, followed by `padding` zero bytes.
bool readWholeFile(const char* path, std::vector<uint8_t>& out, size_t padding = 0);
bool readSuffixArray(const char* path, std::vector<int32_t>& out);
bool writeWholeFile(const char* path, const void* data, size_t size);
bool readBwtFile(const char* path, std::vector<uint8_t>& out, int32_t& primary);
bool writeBwtFile(const char* path, const uint8_t* bwt, int32_t n, int32_t primary);

// Strict .npy reader for the oss_input directory.
//
// Replaces the CNPY loader the reference uses (External/CNPY/cnpy.cpp:178-196,
// called from src/SpikeSorter/OnlineSpikesV2.cu:580-598 and
// src/SpikeSorter/SorterHelpers.cpp:24-115).  Unlike CNPY it honours
// fortran_order (transposing to C order on load), checks dtype and rank
// against what the caller expects, and reports errors as exceptions that the
// C-ABI layer turns into status codes.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace lss {

struct NpyArray {
    std::vector<int64_t> shape;
    char kind = 0;        // 'f', 'i', 'u', 'b'
    int itemsize = 0;
    std::vector<uint8_t> bytes;   // C-order data
    int64_t numel() const {
        int64_t n = 1;
        for (auto s : shape) n *= s;
        return n;
    }
    // Convert to the requested element type (float/double/int32/int64) with range checks.
    template <class T> std::vector<T> as() const;
};

inline std::string npy_header_value(const std::string& h, const std::string& key) {
    auto p = h.find("'" + key + "'");
    if (p == std::string::npos) throw std::runtime_error("npy header lacks key " + key);
    p = h.find(':', p);
    if (p == std::string::npos) throw std::runtime_error("bad npy header");
    ++p;
    while (p < h.size() && h[p] == ' ') ++p;
    if (h[p] == '\'') {
        auto q = h.find('\'', p + 1);
        return h.substr(p + 1, q - p - 1);
    }
    if (h[p] == '(') {
        auto q = h.find(')', p);
        return h.substr(p, q - p + 1);
    }
    auto q = h.find_first_of(",}", p);
    return h.substr(p, q - p);
}

inline NpyArray npy_load(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open " + path);
    struct Closer { FILE* f; ~Closer() { std::fclose(f); } } closer{f};
    unsigned char magic[8];
    if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, "\x93NUMPY", 6) != 0)
        throw std::runtime_error(path + ": not a .npy file");
    uint32_t hlen = 0;
    if (magic[6] == 1) {
        uint16_t h16;
        if (std::fread(&h16, 2, 1, f) != 1) throw std::runtime_error(path + ": truncated");
        hlen = h16;
    } else if (magic[6] == 2 || magic[6] == 3) {
        if (std::fread(&hlen, 4, 1, f) != 1) throw std::runtime_error(path + ": truncated");
    } else {
        throw std::runtime_error(path + ": unsupported .npy version");
    }
    std::string h(hlen, '\0');
    if (std::fread(&h[0], 1, hlen, f) != hlen) throw std::runtime_error(path + ": truncated header");
    NpyArray a;
    std::string descr = npy_header_value(h, "descr");
    if (descr.size() < 3) throw std::runtime_error(path + ": bad descr " + descr);
    if (descr[0] == '>') throw std::runtime_error(path + ": big-endian data not supported");
    if (descr[0] == '|' && descr[1] != 'b' && descr[1] != 'i' && descr[1] != 'u')
        throw std::runtime_error(path + ": unsupported dtype " + descr);
    a.kind = descr[1];
    a.itemsize = std::stoi(descr.substr(2));
    if (a.kind != 'f' && a.kind != 'i' && a.kind != 'u' && a.kind != 'b')
        throw std::runtime_error(path + ": unsupported dtype " + descr + " (object arrays are refused)");
    bool fortran = npy_header_value(h, "fortran_order").find("True") != std::string::npos;
    std::string sh = npy_header_value(h, "shape");
    for (size_t i = 1; i < sh.size();) {
        while (i < sh.size() && (sh[i] == ' ' || sh[i] == ',')) ++i;
        if (i >= sh.size() || sh[i] == ')') break;
        size_t j = i;
        while (j < sh.size() && sh[j] >= '0' && sh[j] <= '9') ++j;
        if (j == i) throw std::runtime_error(path + ": bad shape " + sh);
        a.shape.push_back(std::stoll(sh.substr(i, j - i)));
        i = j;
    }
    const size_t nbytes = (size_t)a.numel() * a.itemsize;
    std::vector<uint8_t> raw(nbytes);
    if (nbytes && std::fread(raw.data(), 1, nbytes, f) != nbytes)
        throw std::runtime_error(path + ": truncated data");
    if (fortran && a.shape.size() > 1) {
        // transpose column-major bytes to row-major
        const int nd = (int)a.shape.size();
        std::vector<int64_t> cstride(nd), fstride(nd);
        cstride[nd - 1] = 1;
        for (int d = nd - 2; d >= 0; --d) cstride[d] = cstride[d + 1] * a.shape[d + 1];
        fstride[0] = 1;
        for (int d = 1; d < nd; ++d) fstride[d] = fstride[d - 1] * a.shape[d - 1];
        a.bytes.resize(nbytes);
        const int64_t n = a.numel();
        std::vector<int64_t> idx(nd, 0);
        for (int64_t lin = 0; lin < n; ++lin) {
            int64_t fo = 0;
            for (int d = 0; d < nd; ++d) fo += idx[d] * fstride[d];
            std::memcpy(&a.bytes[(size_t)lin * a.itemsize], &raw[(size_t)fo * a.itemsize], a.itemsize);
            for (int d = nd - 1; d >= 0; --d) {
                if (++idx[d] < a.shape[d]) break;
                idx[d] = 0;
            }
        }
    } else {
        a.bytes.swap(raw);
    }
    return a;
}

template <class T> std::vector<T> NpyArray::as() const {
    const int64_t n = numel();
    std::vector<T> out((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        const uint8_t* p = &bytes[(size_t)i * itemsize];
        double v = 0;
        if (kind == 'f' && itemsize == 4) { float x; std::memcpy(&x, p, 4); v = x; }
        else if (kind == 'f' && itemsize == 8) { double x; std::memcpy(&x, p, 8); v = x; }
        else if (kind == 'i' && itemsize == 8) { int64_t x; std::memcpy(&x, p, 8); v = (double)x; }
        else if (kind == 'i' && itemsize == 4) { int32_t x; std::memcpy(&x, p, 4); v = x; }
        else if (kind == 'i' && itemsize == 2) { int16_t x; std::memcpy(&x, p, 2); v = x; }
        else if (kind == 'u' && itemsize == 8) { uint64_t x; std::memcpy(&x, p, 8); v = (double)x; }
        else if (kind == 'u' && itemsize == 4) { uint32_t x; std::memcpy(&x, p, 4); v = x; }
        else if ((kind == 'u' || kind == 'b') && itemsize == 1) { v = *p; }
        else if (kind == 'i' && itemsize == 1) { v = (int8_t)*p; }
        else throw std::runtime_error("unsupported npy element type");
        out[(size_t)i] = (T)v;
    }
    return out;
}

}  // namespace lss

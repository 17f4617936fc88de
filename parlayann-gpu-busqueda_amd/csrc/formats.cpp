// formats.cpp -- readers for the search path's on-disk formats (see formats.h).
#include "formats.h"

#include <cstdio>
#include <algorithm>
#include <cstring>

namespace pann {
namespace {

constexpr int kErrIo = 2, kErrFormat = 3;

Status fail(int code, const std::string& m) {
    Status s;
    s.code = code;
    s.msg = m;
    return s;
}

bool readAll(const std::string& path, std::vector<char>& buf) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) return false;
    if (fseeko(f, 0, SEEK_END) != 0) {
        fclose(f);
        return false;
    }
    off_t size = ftello(f);
    fseeko(f, 0, SEEK_SET);
    buf.resize((size_t)size);
    size_t got = size > 0 ? fread(buf.data(), 1, (size_t)size, f) : 0;
    fclose(f);
    return got == (size_t)size;
}

struct Cursor {
    const std::vector<char>& buf;
    size_t pos = 0;
    bool bad = false;
    template <typename T>
    T get() {
        T x{};
        if (pos + sizeof(T) > buf.size()) {
            bad = true;
            return x;
        }
        memcpy(&x, buf.data() + pos, sizeof(T));
        pos += sizeof(T);
        return x;
    }
};

}  // namespace

Status readI8bin(const std::string& path, HostVectors& out) {
    std::vector<char> buf;
    if (!readAll(path, buf)) return fail(kErrIo, "cannot read " + path);
    if (buf.size() < 8) return fail(kErrFormat, path + ": shorter than the 8-byte header");
    uint32_t hdr[2];
    memcpy(hdr, buf.data(), 8);
    const uint64_t expect = 8 + (uint64_t)hdr[0] * hdr[1];
    if (buf.size() != expect)
        return fail(kErrFormat, path + ": size " + std::to_string(buf.size()) +
                                    " but the header says " + std::to_string(expect));
    out.n = hdr[0];
    out.dim = hdr[1];
    out.data.assign(reinterpret_cast<const int8_t*>(buf.data()) + 8,
                    reinterpret_cast<const int8_t*>(buf.data()) + buf.size());
    return {};
}

Status checkGraph(uint32_t n, uint32_t maxDeg, const uint32_t* degrees, const uint32_t* edges,
                  uint64_t* totalEdges) {
    uint64_t total = 0;
    for (uint32_t i = 0; i < n; i++) {
        if (degrees[i] > maxDeg)
            return fail(kErrFormat, "node " + std::to_string(i) + " has degree " +
                                        std::to_string(degrees[i]) + " > max degree " +
                                        std::to_string(maxDeg));
        total += degrees[i];
    }
    for (uint64_t e = 0; e < total; e++)
        if (edges[e] >= n)
            return fail(kErrFormat, "edge " + std::to_string(e) + " points to " +
                                        std::to_string(edges[e]) + " (n=" + std::to_string(n) +
                                        ")");
    if (totalEdges) *totalEdges = total;
    return {};
}

Status readParlayGraph(const std::string& path, HostGraph& out) {
    std::vector<char> buf;
    if (!readAll(path, buf)) return fail(kErrIo, "cannot read " + path);
    if (buf.size() < 8) return fail(kErrFormat, path + ": shorter than the 8-byte header");
    uint32_t hdr[2];
    memcpy(hdr, buf.data(), 8);
    const uint32_t n = hdr[0];
    if (buf.size() < 8 + 4ull * n) return fail(kErrFormat, path + ": truncated degree array");
    out.n = n;
    out.maxDeg = hdr[1];
    out.degrees.resize(n);
    memcpy(out.degrees.data(), buf.data() + 8, 4ull * n);
    out.offsets.assign((size_t)n + 1, 0);
    for (uint32_t i = 0; i < n; i++) {
        if (out.degrees[i] > out.maxDeg)
            return fail(kErrFormat, path + ": node " + std::to_string(i) + " degree " +
                                        std::to_string(out.degrees[i]) + " > max degree " +
                                        std::to_string(out.maxDeg));
        out.offsets[i + 1] = out.offsets[i] + out.degrees[i];
    }
    const uint64_t total = out.offsets[n];
    const uint64_t expect = 8 + 4ull * n + 4ull * total;
    if (buf.size() != expect)
        return fail(kErrFormat, path + ": size " + std::to_string(buf.size()) + " but should be " +
                                    std::to_string(expect));
    out.edges.resize(total);
    memcpy(out.edges.data(), buf.data() + 8 + 4ull * n, 4ull * total);
    for (uint64_t e = 0; e < total; e++)
        if (out.edges[e] >= n)
            return fail(kErrFormat, path + ": edge " + std::to_string(e) + " points to " +
                                        std::to_string(out.edges[e]) + " (n=" +
                                        std::to_string(n) + ")");
    return {};
}

Status readHnsw(const std::string& path, HostHnsw& h) {
    std::vector<char> buf;
    if (!readAll(path, buf)) return fail(kErrIo, "cannot read " + path);
    Cursor c{buf};
    char magic[4];
    for (int i = 0; i < 4; i++) magic[i] = c.get<char>();
    if (c.bad || memcmp(magic, "HNSW", 4) != 0) return fail(kErrFormat, path + ": not an HNSW model");
    const uint32_t version = c.get<uint32_t>();
    if (c.bad || version != 3)
        return fail(kErrFormat, path + ": version " + std::to_string(version) + ", expected 3");
    c.get<uint64_t>();  // code_U
    c.get<uint64_t>();  // size_node
    h.dim = c.get<uint32_t>();
    h.ml = c.get<float>();
    h.m = c.get<uint32_t>();
    h.efc = c.get<uint32_t>();
    h.alpha = c.get<float>();
    h.n = c.get<uint32_t>();
    if (c.bad) return fail(kErrFormat, path + ": truncated header");
    if ((uint64_t)h.n * 8 > buf.size()) return fail(kErrFormat, path + ": truncated node table");
    h.level.resize(h.n);
    h.nodePoint.resize(h.n);
    for (uint32_t i = 0; i < h.n; i++) {
        h.level[i] = c.get<uint32_t>();
        h.nodePoint[i] = c.get<uint32_t>();
    }
    if (c.bad) return fail(kErrFormat, path + ": truncated node table");
    h.firstList.resize(h.n);
    h.listOffset.clear();
    h.listSize.clear();
    h.ids.clear();
    h.maxDeg0 = h.maxDegUpper = 0;
    for (uint32_t i = 0; i < h.n; i++) {
        h.firstList[i] = h.listOffset.size();
        for (uint32_t l = 0; l <= h.level[i]; l++) {
            const uint64_t cnt = c.get<uint64_t>();
            if (c.bad || cnt > (buf.size() - c.pos) / 4)
                return fail(kErrFormat, path + ": truncated neighbour list");
            h.listOffset.push_back(h.ids.size());
            h.listSize.push_back((uint32_t)cnt);
            for (uint64_t j = 0; j < cnt; j++) {
                const uint32_t v = c.get<uint32_t>();
                if (v >= h.n)
                    return fail(kErrFormat, path + ": neighbour " + std::to_string(v) +
                                                " out of range (n=" + std::to_string(h.n) + ")");
                h.ids.push_back(v);
            }
            if (l == 0) h.maxDeg0 = std::max<uint32_t>(h.maxDeg0, (uint32_t)cnt);
            else h.maxDegUpper = std::max<uint32_t>(h.maxDegUpper, (uint32_t)cnt);
        }
    }
    const uint64_t ne = c.get<uint64_t>();
    if (c.bad || ne > (buf.size() - c.pos) / 4)
        return fail(kErrFormat, path + ": truncated entry list");
    h.entries.resize(ne);
    for (uint64_t i = 0; i < ne; i++) h.entries[i] = c.get<uint32_t>();
    if (c.pos != buf.size())
        return fail(kErrFormat, path + ": " + std::to_string(buf.size() - c.pos) +
                                    " trailing bytes, unexpected format");
    if (h.entries.empty() || h.entries.size() > 8)
        return fail(kErrFormat, path + ": " + std::to_string(h.entries.size()) +
                                    " entry points (1..8 supported)");
    for (uint32_t e : h.entries)
        if (e >= h.n) return fail(kErrFormat, path + ": entry point out of range");
    return {};
}

}  // namespace pann

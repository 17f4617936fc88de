// tools.cpp -- error reporting and file writers of the host tools library.
#include <cstdio>
#include <string>

#include "../../include/pann_tools.h"
#include "tools_common.h"

namespace pann {
namespace tools {

namespace {
thread_local std::string g_err;
}

int setError(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

// ParlayANN Graph::save layout: [n][maxDeg][degrees n][edges]   cuda/lectores.h:17-27
int writeParlayGraph(const char* path, uint32_t n, uint32_t maxDeg, const uint32_t* degrees,
                     const uint32_t* dense, uint32_t stride) {
    FILE* f = fopen(path, "wb");
    if (!f) return setError(2, std::string("cannot write ") + path);
    const uint32_t hdr[2] = {n, maxDeg};
    bool ok = fwrite(hdr, 4, 2, f) == 2;
    ok = ok && fwrite(degrees, 4, n, f) == n;
    for (uint32_t i = 0; ok && i < n; i++)
        ok = fwrite(dense + (size_t)i * stride, 4, degrees[i], f) == degrees[i];
    ok = (fclose(f) == 0) && ok;
    if (!ok) return setError(2, std::string("short write to ") + path);
    return 0;
}

size_t bridgeFromStart(const int8_t* base, uint32_t n, uint32_t dim,
                       std::vector<std::vector<uint32_t>>& adj, int nthreads) {
    std::vector<uint8_t> reach(n, 0);
    std::vector<uint32_t> stack;
    auto flood = [&](uint32_t s) {
        reach[s] = 1;
        stack.push_back(s);
        while (!stack.empty()) {
            const uint32_t u = stack.back();
            stack.pop_back();
            for (uint32_t a : adj[u])
                if (!reach[a]) {
                    reach[a] = 1;
                    stack.push_back(a);
                }
        }
    };
    flood(0);
    size_t added = 0;
    const int T = std::max(1, nthreads);
    for (uint32_t x = 0; x < n; x++) {
        if (reach[x]) continue;
        std::vector<Nbr> best(T, Nbr{INT32_MAX, 0xFFFFFFFFu});
        const int8_t* xv = base + (size_t)x * dim;
        parallelFor(n, T, [&](size_t y, int t) {
            if (!reach[y]) return;
            const Nbr c{distI8(xv, base + y * dim, dim), (uint32_t)y};
            if (c < best[t]) best[t] = c;
        }, 4096);
        const Nbr b = *std::min_element(best.begin(), best.end());
        adj[b.id].push_back(x);
        added++;
        flood(x);
    }
    return added;
}

int writeGraphLists(const char* path, const std::vector<std::vector<uint32_t>>& adj) {
    const uint32_t n = (uint32_t)adj.size();
    uint32_t maxDeg = 0;
    for (const auto& a : adj) maxDeg = std::max<uint32_t>(maxDeg, (uint32_t)a.size());
    FILE* f = fopen(path, "wb");
    if (!f) return setError(2, std::string("cannot write ") + path);
    const uint32_t hdr[2] = {n, maxDeg};
    bool ok = fwrite(hdr, 4, 2, f) == 2;
    for (uint32_t i = 0; ok && i < n; i++) {
        const uint32_t d = (uint32_t)adj[i].size();
        ok = fwrite(&d, 4, 1, f) == 1;
    }
    for (uint32_t i = 0; ok && i < n; i++)
        ok = fwrite(adj[i].data(), 4, adj[i].size(), f) == adj[i].size();
    ok = (fclose(f) == 0) && ok;
    if (!ok) return setError(2, std::string("short write to ") + path);
    return 0;
}

}  // namespace tools
}  // namespace pann

extern "C" {

const char* pann_tools_last_error(void) { return pann::tools::g_err.c_str(); }

int pann_write_graph(const char* path, uint32_t n, uint32_t max_degree, const uint32_t* degrees,
                     const uint32_t* edges_dense, uint32_t stride) {
    if (!path || !degrees || (!edges_dense && n)) return pann::tools::setError(1, "NULL argument");
    for (uint32_t i = 0; i < n; i++)
        if (degrees[i] > max_degree || degrees[i] > stride)
            return pann::tools::setError(1, "degree above max_degree/stride");
    return pann::tools::writeParlayGraph(path, n, max_degree, degrees, edges_dense, stride);
}

}  // extern "C"

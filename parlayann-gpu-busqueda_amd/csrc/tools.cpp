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

int writeGraphLists(const char* path, const std::vector<std::vector<uint32_t>>& adj) {
    const int r = writeGraphListsFile(path, adj);
    if (r == 1) return setError(2, std::string("cannot write ") + path);
    if (r == 2) return setError(2, std::string("short write to ") + path);
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

// host_io.cpp -- operator and stencil-graph file I/O of the host layer:
// the points file (.points), the stencil graph (.graph, ragged CSR) and the
// Matrix Market dump of an assembled operator. The formats are the
// reference's (rbf_io.h:194-232, 630-707, 771-823): the same header lines,
// the same validation, numbers written as the shortest decimal text that
// reads back exactly (std::to_chars), indices 1-based in Matrix Market and
// 0-based in .graph.

#include <charconv>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <string>
#include <string_view>
#include <system_error>
#include <vector>

#include "rbf_assembly.h"

namespace {

thread_local std::string g_last_error;

int32_t fail(int32_t status, const std::string& fname, const std::string& what) {
    g_last_error = fname + ": " + what;
    return status;
}

bool is_ws(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n'; }

// Tokens of one line; a number must end at whitespace or the end of the
// line (so "1.5" is not the integer 1).
struct Cursor {
    const char* p;
    const char* end;
    explicit Cursor(std::string_view s) : p(s.data()), end(s.data() + s.size()) {}
    void skip() {
        while (p != end && is_ws(*p)) ++p;
    }
    bool at_end() {
        skip();
        return p == end;
    }
    template <class V>
    bool next(V& v) {
        skip();
        if (p == end) return false;
        const auto r = std::from_chars(p, end, v);
        if (r.ec != std::errc() || (r.ptr != end && !is_ws(*r.ptr))) return false;
        p = r.ptr;
        return true;
    }
};

}  // namespace

namespace rbf_amd_internal {
// declared for the kernel translation units in assemble_common.h
void set_last_error(const char* what) { g_last_error = what; }
}  // namespace rbf_amd_internal

extern "C" {

RBFA_API const char* rbfa_last_error(void) { return g_last_error.c_str(); }

RBFA_API int32_t rbfa_read_points(const char* path, int32_t cap, int32_t* n, double* x, double* y) {
    if (!path || !n || cap < 0 || (cap > 0 && (!x || !y))) return RBFA_ERR_ARG;
    std::ifstream in(path);
    if (!in) return fail(RBFA_ERR_IO, path, "cannot open for reading");
    std::string line;
    std::getline(in, line);
    int64_t count = 0;
    if (Cursor c{line}; !c.next(count) || !c.at_end() || count < 0 || count > INT32_MAX)
        return fail(RBFA_ERR_FORMAT, path, "missing or malformed point count");
    *n = static_cast<int32_t>(count);
    if (cap == 0) return RBFA_OK;
    if (cap < count) return fail(RBFA_ERR_ARG, path, "output arrays hold fewer than n points");
    for (int64_t i = 0; i < count; ++i) {
        if (!std::getline(in, line))
            return fail(RBFA_ERR_FORMAT, path,
                        "expected " + std::to_string(count) + " points, found " + std::to_string(i));
        Cursor c{line};
        double px, py;
        if (!c.next(px) || !c.next(py) || !c.at_end())
            return fail(RBFA_ERR_FORMAT, path,
                        "line " + std::to_string(i + 2) + ": expected two coordinates");
        x[i] = px;
        y[i] = py;
    }
    return RBFA_OK;
}

RBFA_API int32_t rbfa_read_graph(const char* path, int32_t k, int32_t* n, int64_t* nnz,
                                 int32_t* ia, int32_t* ja, int64_t cap) {
    if (!path || !n || !nnz || k < 0 || ((ia == nullptr) != (ja == nullptr))) return RBFA_ERR_ARG;
    std::ifstream in(path);
    if (!in) return fail(RBFA_ERR_IO, path, "cannot open for reading");
    std::string line;
    std::getline(in, line);
    int64_t rows = 0, entries = 0;
    if (Cursor c{line}; !c.next(rows) || !c.next(entries) || !c.at_end() || rows < 0 ||
                        entries < 0 || rows > INT32_MAX || entries > INT32_MAX)
        return fail(RBFA_ERR_FORMAT, path, "malformed header, expected: n nnz");
    if (k > 0 && entries != rows * k)
        return fail(RBFA_ERR_FORMAT, path,
                    "header says " + std::to_string(rows) + " rows and " +
                        std::to_string(entries) + " entries, inconsistent with k=" +
                        std::to_string(k));
    *n = static_cast<int32_t>(rows);
    *nnz = entries;
    if (!ia) return RBFA_OK;  // sizes only
    if (cap < entries) return fail(RBFA_ERR_ARG, path, "ja holds fewer than nnz entries");
    std::vector<int64_t> last_row(static_cast<size_t>(rows), -1);
    int64_t e = 0;
    ia[0] = 0;
    for (int64_t i = 0; i < rows; ++i) {
        if (!std::getline(in, line))
            return fail(RBFA_ERR_FORMAT, path,
                        "header says " + std::to_string(rows) + " rows, found " + std::to_string(i));
        Cursor c{line};
        int64_t len = 0;
        while (!c.at_end()) {
            int64_t j;
            if (!c.next(j))
                return fail(RBFA_ERR_FORMAT, path, "row " + std::to_string(i) + ": bad index");
            if (j < 0 || j >= rows)
                return fail(RBFA_ERR_FORMAT, path,
                            "row " + std::to_string(i) + ": index " + std::to_string(j) +
                                " outside [0, " + std::to_string(rows) +
                                "); indices are 0-based");
            if (last_row[j] == i)
                return fail(RBFA_ERR_FORMAT, path,
                            "row " + std::to_string(i) + ": index " + std::to_string(j) +
                                " listed twice");
            last_row[j] = i;
            if (e >= entries)
                return fail(RBFA_ERR_FORMAT, path,
                            "header says " + std::to_string(entries) + " entries, found more");
            ja[e++] = static_cast<int32_t>(j);
            ++len;
        }
        if (len == 0) return fail(RBFA_ERR_FORMAT, path, "row " + std::to_string(i) + " has no entries");
        if (k > 0 && len != k)
            return fail(RBFA_ERR_FORMAT, path,
                        "row " + std::to_string(i) + " has " + std::to_string(len) +
                            " entries, expected k=" + std::to_string(k));
        ia[i + 1] = static_cast<int32_t>(e);
    }
    if (e != entries)
        return fail(RBFA_ERR_FORMAT, path,
                    "header says " + std::to_string(entries) + " entries, found " +
                        std::to_string(e));
    in >> std::ws;
    if (!in.eof())
        return fail(RBFA_ERR_FORMAT, path,
                    "unexpected data after row " + std::to_string(rows - 1));
    return RBFA_OK;
}

RBFA_API int32_t rbfa_write_matrix_market(const char* path, int64_t rows, int64_t cols,
                                          const int32_t* ia, const int32_t* ja, const double* a,
                                          int32_t csr_base) {
    if (!path || rows < 0 || cols < 0 || !ia || (rows > 0 && !ja) ||
        (csr_base != 0 && csr_base != 1) || ia[0] != csr_base)
        return RBFA_ERR_ARG;
    for (int64_t i = 0; i < rows; ++i) {
        if (ia[i + 1] < ia[i]) return fail(RBFA_ERR_ARG, path, "ia is not non-decreasing");
        for (int64_t e = ia[i] - csr_base; e < ia[i + 1] - csr_base; ++e)
            if (ja[e] < csr_base || ja[e] - csr_base >= cols)
                return fail(RBFA_ERR_ARG, path, "ja column out of range");
    }
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return fail(RBFA_ERR_IO, path, "cannot open for writing");
    std::string buf;
    buf.reserve(1 << 20);
    buf += "%%MatrixMarket matrix coordinate ";
    buf += a ? "real" : "pattern";
    buf += " general\n";
    buf += std::to_string(rows) + ' ' + std::to_string(cols) + ' ' +
           std::to_string(static_cast<int64_t>(ia[rows]) - ia[0]) + '\n';
    char num[64];
    bool ok = true;
    for (int64_t i = 0; i < rows && ok; ++i) {
        for (int64_t e = ia[i] - csr_base; e < ia[i + 1] - csr_base; ++e) {
            auto r = std::to_chars(num, num + sizeof num, i + 1);
            buf.append(num, r.ptr);
            buf += ' ';
            r = std::to_chars(num, num + sizeof num, static_cast<int64_t>(ja[e]) - csr_base + 1);
            buf.append(num, r.ptr);
            if (a) {
                buf += ' ';
                r = std::to_chars(num, num + sizeof num, a[e]);  // shortest exact text
                buf.append(num, r.ptr);
            }
            buf += '\n';
        }
        if (buf.size() > (1u << 20)) {
            ok = std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
            buf.clear();
        }
    }
    if (ok && !buf.empty()) ok = std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
    if (std::fclose(f) != 0) ok = false;
    return ok ? RBFA_OK : fail(RBFA_ERR_IO, path, "write failed");
}

}  // extern "C"

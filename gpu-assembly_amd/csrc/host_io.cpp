// host_io.cpp -- operator and stencil-graph file I/O of the host layer:
// the points file (.points), the stencil graph (.graph, ragged CSR) and the
// Matrix Market dump of an assembled operator. The formats and the rules a
// file must satisfy are the reference's (rbf_io.h:194-232, 630-707,
// 771-823; docs/file_formats.md); numbers are written as the shortest
// decimal text that reads back exactly (std::to_chars), indices 1-based in
// Matrix Market and 0-based in .graph.
//
// The readers load the whole file, split it into lines and convert each
// line's tokens in one go; the graph reader then checks the structure of
// the rows it stored (lengths, duplicates by sorting a copy of each row)
// before it checks the totals and the end of the file. Every failure
// leaves a "<file>: <what>" text for rbfa_last_error.

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <sstream>
#include <string>
#include <string_view>
#include <system_error>
#include <vector>

#include "rbf_assembly.h"

namespace {

thread_local std::string g_last_error;

// Records "<file>: <parts...>" as the calling thread's last error and
// returns the status to hand back.
template <class... Parts>
int32_t report(int32_t status, const char* file, const Parts&... parts) {
    std::ostringstream os;
    os << file << ": ";
    (os << ... << parts);
    g_last_error = os.str();
    return status;
}

// The whole file as text; false if it cannot be opened or read.
bool load_text(const char* path, std::string& text) {
    std::FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    text.clear();
    char chunk[1 << 16];
    size_t got;
    while ((got = std::fread(chunk, 1, sizeof chunk, f)) > 0) text.append(chunk, got);
    const bool ok = !std::ferror(f);
    std::fclose(f);
    return ok;
}

bool blank(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n' || c == '\v' || c == '\f'; }

// Lines of a text, numbered from 1; a '\r' before the '\n' belongs to the
// line break.
class LineReader {
  public:
    explicit LineReader(std::string_view text) : rest_(text) {}
    bool next(std::string_view& line) {
        if (rest_.empty()) return false;
        const size_t cut = rest_.find('\n');
        line = rest_.substr(0, cut);
        rest_ = cut == std::string_view::npos ? std::string_view{} : rest_.substr(cut + 1);
        if (!line.empty() && line.back() == '\r') line.remove_suffix(1);
        ++number_;
        return true;
    }
    int64_t number() const { return number_; }
    // true when only white space is left
    bool spent() const { return std::all_of(rest_.begin(), rest_.end(), blank); }

  private:
    std::string_view rest_;
    int64_t number_ = 0;
};

// Converts every white-space separated token of line to V, appending to
// out. A token that is not entirely a V stops the scan: it is returned in
// bad (empty when the whole line converted).
template <class V>
void tokens_as(std::string_view line, std::vector<V>& out, std::string_view& bad) {
    bad = {};
    const char* p = line.data();
    const char* end = p + line.size();
    while (true) {
        while (p != end && blank(*p)) ++p;
        if (p == end) return;
        const char* q = p;
        while (q != end && !blank(*q)) ++q;
        V v{};
        const auto r = std::from_chars(p, q, v);
        if (r.ec != std::errc() || r.ptr != q) {
            bad = std::string_view(p, static_cast<size_t>(q - p));
            return;
        }
        out.push_back(v);
        p = q;
    }
}

}  // namespace

namespace rbf_amd_internal {
// declared for the kernel translation units in assemble_common.h
void set_last_error(const char* what) { g_last_error = what; }
}  // namespace rbf_amd_internal

extern "C" {

RBFA_API const char* rbfa_last_error(void) { return g_last_error.c_str(); }

RBFA_API int32_t rbfa_read_points(const char* path, int32_t cap, int32_t* n, double* x, double* y) {
    if (!path || !n || cap < 0 || (cap > 0 && (!x || !y))) return RBFA_ERR_ARG;
    std::string text;
    if (!load_text(path, text)) return report(RBFA_ERR_IO, path, "cannot open or read the file");
    LineReader lines(text);
    std::string_view line, bad;
    std::vector<int64_t> head;
    if (lines.next(line)) tokens_as(line, head, bad);
    if (head.size() != 1 || !bad.empty() || head[0] < 0 || head[0] > INT32_MAX)
        return report(RBFA_ERR_FORMAT, path, "line 1 must hold the node count and nothing else");
    const int64_t count = head[0];
    *n = static_cast<int32_t>(count);
    if (cap == 0) return RBFA_OK;  // the count only
    if (cap < count)
        return report(RBFA_ERR_ARG, path, "room for ", cap, " nodes, the file has ", count);
    std::vector<double> xy;
    for (int64_t i = 0; i < count; ++i) {
        if (!lines.next(line))
            return report(RBFA_ERR_FORMAT, path, "the file ends after ", i, " of ", count,
                          " coordinate pairs");
        xy.clear();
        tokens_as(line, xy, bad);
        if (xy.size() != 2 || !bad.empty())
            return report(RBFA_ERR_FORMAT, path, "line ", lines.number(),
                          ": a node takes exactly two coordinates");
        if (!std::isfinite(xy[0]) || !std::isfinite(xy[1]))
            return report(RBFA_ERR_FORMAT, path, "line ", lines.number(),
                          ": coordinates must be finite");
        x[i] = xy[0];
        y[i] = xy[1];
    }
    return RBFA_OK;  // lines after the last pair are not part of the format
}

RBFA_API int32_t rbfa_read_graph(const char* path, int32_t k, int32_t* n, int64_t* nnz,
                                 int32_t* ia, int32_t* ja, int64_t cap) {
    if (!path || !n || !nnz || k < 0 || ((ia == nullptr) != (ja == nullptr))) return RBFA_ERR_ARG;
    std::string text;
    if (!load_text(path, text)) return report(RBFA_ERR_IO, path, "cannot open or read the file");
    LineReader lines(text);
    std::string_view line, bad;
    std::vector<int64_t> head;
    if (lines.next(line)) tokens_as(line, head, bad);
    if (head.size() != 2 || !bad.empty() || head[0] < 0 || head[1] < 0 || head[0] > INT32_MAX ||
        head[1] > INT32_MAX)
        return report(RBFA_ERR_FORMAT, path,
                      "line 1 must hold two counts, the nodes and the stencil entries");
    const int64_t rows = head[0], entries = head[1];
    if (k > 0 && entries != rows * k)
        return report(RBFA_ERR_FORMAT, path, rows, " stencils of fixed size k = ", k, " need ",
                      rows * k, " entries; line 1 declares ", entries);
    *n = static_cast<int32_t>(rows);
    *nnz = entries;
    if (!ia) return RBFA_OK;  // the sizes only
    if (cap < entries)
        return report(RBFA_ERR_ARG, path, "room for ", cap, " entries, the file declares ", entries);

    // 1. tokens: one row per line, every index a node number
    std::vector<int64_t> row;
    int64_t stored = 0;
    ia[0] = 0;
    for (int64_t r = 0; r < rows; ++r) {
        if (!lines.next(line))
            return report(RBFA_ERR_FORMAT, path, "the file ends after ", r, " of ", rows,
                          " stencil rows");
        row.clear();
        tokens_as(line, row, bad);
        if (!bad.empty())
            return report(RBFA_ERR_FORMAT, path, "stencil row ", r, ": '", bad,
                          "' is not an integer node index");
        for (const int64_t j : row)
            if (j < 0 || j >= rows)
                return report(RBFA_ERR_FORMAT, path, "stencil row ", r, ": node index ", j,
                              " lies outside 0..", rows - 1, " (nodes are numbered from 0)");
        if (stored + static_cast<int64_t>(row.size()) > entries)
            return report(RBFA_ERR_FORMAT, path, "the rows hold more than the ", entries,
                          " entries line 1 declares");
        for (const int64_t j : row) ja[stored++] = static_cast<int32_t>(j);
        ia[r + 1] = static_cast<int32_t>(stored);
    }
    // 2. structure of the stored rows
    std::vector<int32_t> sorted;
    for (int64_t r = 0; r < rows; ++r) {
        const int32_t* first = ja + ia[r];
        const int32_t* last = ja + ia[r + 1];
        const int64_t len = last - first;
        if (len == 0) return report(RBFA_ERR_FORMAT, path, "stencil row ", r, " is empty");
        if (k > 0 && len != k)
            return report(RBFA_ERR_FORMAT, path, "stencil row ", r, " holds ", len,
                          " nodes; fixed-size stencils hold ", k);
        sorted.assign(first, last);
        std::sort(sorted.begin(), sorted.end());
        const auto twice = std::adjacent_find(sorted.begin(), sorted.end());
        if (twice != sorted.end())
            return report(RBFA_ERR_FORMAT, path, "stencil row ", r, " names node ", *twice,
                          " twice");
    }
    // 3. totals and the end of the file
    if (stored != entries)
        return report(RBFA_ERR_FORMAT, path, "the rows hold ", stored, " entries; line 1 declares ",
                      entries);
    if (!lines.spent())
        return report(RBFA_ERR_FORMAT, path, "text follows the last of the ", rows,
                      " stencil rows");
    return RBFA_OK;
}

RBFA_API int32_t rbfa_write_matrix_market(const char* path, int64_t rows, int64_t cols,
                                          const int32_t* ia, const int32_t* ja, const double* a,
                                          int32_t csr_base) {
    if (!path || rows < 0 || cols < 0 || !ia || (rows > 0 && !ja) ||
        (csr_base != 0 && csr_base != 1) || ia[0] != csr_base)
        return RBFA_ERR_ARG;
    for (int64_t i = 0; i < rows; ++i) {
        if (ia[i + 1] < ia[i]) return report(RBFA_ERR_ARG, path, "ia is not non-decreasing");
        for (int64_t e = ia[i] - csr_base; e < ia[i + 1] - csr_base; ++e)
            if (ja[e] < csr_base || ja[e] - csr_base >= cols)
                return report(RBFA_ERR_ARG, path, "ja column out of range");
    }
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return report(RBFA_ERR_IO, path, "cannot open for writing");
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
    return ok ? RBFA_OK : report(RBFA_ERR_IO, path, "write failed");
}

}  // extern "C"

// Matrix Market (.mtx) and binary (.mtxbin) edge-list files on the host.
//
// Replaces Matrix::init_from_mtx's readers (src/backend/matrix/read_file.hpp:
// read_mtx_file_pipelined 52-170, binary_read_mtx_file 378-555, dispatch by
// suffix 556-590) and the writers EdgeListContainer::save_as_mtx
// (src/helpers/graph_generation/edges_container.hpp:33-88) and
// benchmark/scripts/mtx_api.py:61-90 (the .mtxbin layout).
//
// .mtx: lines starting with '%' before the size line are skipped; the size
// line is "rows cols entries"; each entry line starts with "row col"
// (1-based), further fields (values) are ignored: the reference reads
// pattern data and gives every entry the value EDGE_VAL = 1.
// .mtxbin: three little-endian int64 (rows, cols, entries), then one int64
// pair (row, col), 1-based, per entry.
//
// Entries come back 0-based in file order.  The text reader maps the file
// and parses newline-aligned chunks on all threads (count, prefix, parse), so
// the order does not depend on the thread count.  The reference prints a
// message for ids <= 0 and continues; here any id outside [1, rows] x
// [1, cols], a short file or an unparsable entry line is an error.
#include <algorithm>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <fcntl.h>
#include <omp.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include "../../include/gbk.h"

namespace gbk {
int fail(int code, const char* fmt, ...);
}
using gbk::fail;

namespace {

bool ends_with(const std::string& s, const char* suf) {
    const size_t n = std::strlen(suf);
    return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

// Parses an unsigned decimal at p (after spaces / tabs); false if none.
inline bool parse_i64(const char*& p, const char* end, int64_t* out) {
    while (p < end && (*p == ' ' || *p == '\t')) ++p;
    bool neg = false;
    if (p < end && (*p == '-' || *p == '+')) neg = (*p++ == '-');
    if (p >= end || *p < '0' || *p > '9') return false;
    int64_t v = 0;
    while (p < end && *p >= '0' && *p <= '9') v = v * 10 + (*p++ - '0');
    *out = neg ? -v : v;
    return true;
}

// Decimal digits of v at p; returns the length.
inline int fmt_u64(char* p, uint64_t v) {
    char d[24];
    int k = 0;
    do {
        d[k++] = (char)('0' + v % 10);
        v /= 10;
    } while (v);
    for (int i = 0; i < k; ++i) p[i] = d[k - 1 - i];
    return k;
}

inline bool blank_line(const char* p, const char* e) {
    for (; p < e; ++p)
        if (*p != ' ' && *p != '\t' && *p != '\r') return false;
    return true;
}

// malloc'd 0-based entry arrays (no zero fill of multi-GB outputs)
struct HostCoo {
    int64_t* rows = nullptr;
    int64_t* cols = nullptr;
    int64_t nnz = 0;
    ~HostCoo() {
        std::free(rows);
        std::free(cols);
    }
    int alloc(int64_t k) {
        const size_t bytes = sizeof(int64_t) * (size_t)(k > 0 ? k : 1);
        rows = (int64_t*)std::malloc(bytes);
        cols = (int64_t*)std::malloc(bytes);
        nnz = k;
        if (!rows || !cols) return fail(GBK_OUT_OF_MEMORY, "host allocation of %zu bytes failed", 2 * bytes);
        return GBK_SUCCESS;
    }
};

struct Mapped {
    const char* data = nullptr;
    size_t size = 0;
    int fd = -1;
    ~Mapped() {
        if (data && size) munmap((void*)data, size);
        if (fd >= 0) close(fd);
    }
};

int read_text(const char* path, int64_t* nrows, int64_t* ncols, HostCoo& out) {
    Mapped m;
    m.fd = open(path, O_RDONLY);
    if (m.fd < 0) return fail(GBK_INVALID_VALUE, "Error: Can not open .mtx file (%s)", path);
    struct stat st;
    if (fstat(m.fd, &st) != 0) return fail(GBK_INVALID_VALUE, "Error: Can not open .mtx file (%s)", path);
    m.size = (size_t)st.st_size;
    if (m.size == 0) return fail(GBK_INVALID_VALUE, "Error: unexpected end of graph file! Aborting...");
    void* p = mmap(nullptr, m.size, PROT_READ, MAP_PRIVATE, m.fd, 0);
    if (p == MAP_FAILED) {
        m.size = 0;
        return fail(GBK_INVALID_VALUE, "mmap of %s failed: %s", path, std::strerror(errno));
    }
    m.data = (const char*)p;
    const char* b = m.data;
    const char* end = m.data + m.size;
    // size line: the first line not starting with '%'
    const char* line = b;
    while (true) {
        if (line >= end) return fail(GBK_INVALID_VALUE, "Error: unexpected end of graph file! Aborting...");
        const char* nl = (const char*)std::memchr(line, '\n', end - line);
        const char* le = nl ? nl : end;
        if (line < le && line[0] == '%') {
            line = nl ? nl + 1 : end;
            continue;
        }
        const char* q = line;
        int64_t r = 0, c = 0, z = 0;
        if (!parse_i64(q, le, &r) || !parse_i64(q, le, &c) || !parse_i64(q, le, &z) || r < 0 || c < 0 || z < 0)
            return fail(GBK_INVALID_VALUE, "malformed .mtx size line");
        *nrows = r;
        *ncols = c;
        line = nl ? nl + 1 : end;
        if (out.alloc(z) != GBK_SUCCESS) return GBK_OUT_OF_MEMORY;
        break;
    }
    const int64_t want = out.nnz;
    int64_t* rows = out.rows;
    int64_t* cols = out.cols;
    if (want == 0) return GBK_SUCCESS;
    // newline-aligned chunks of the body
    const int nt = std::max(1, std::min(omp_get_max_threads(), 64));
    const size_t body = (size_t)(end - line);
    std::vector<const char*> cb(nt + 1);
    cb[0] = line;
    cb[nt] = end;
    for (int t = 1; t < nt; ++t) {
        const char* s = line + body * t / nt;
        if (s < cb[t - 1]) s = cb[t - 1];
        const char* nl = (const char*)std::memchr(s, '\n', end - s);
        cb[t] = nl ? nl + 1 : end;
    }
    std::vector<int64_t> cnt(nt + 1, 0);
    int bad = 0;
#pragma omp parallel num_threads(nt) reduction(| : bad)
    {
        const int t = omp_get_thread_num();
        // pass 1: entry lines of the chunk (blank and '%' lines skipped)
        int64_t k = 0;
        for (const char* s = cb[t]; s < cb[t + 1];) {
            const char* nl = (const char*)std::memchr(s, '\n', cb[t + 1] - s);
            const char* le = nl ? nl : cb[t + 1];
            if (!(s < le && s[0] == '%') && !blank_line(s, le)) ++k;
            s = nl ? nl + 1 : cb[t + 1];
        }
        cnt[t + 1] = k;
#pragma omp barrier
#pragma omp single
        for (int i = 0; i < nt; ++i) cnt[i + 1] += cnt[i];
        // pass 2: parse the entries that fall among the first `want`
        int64_t o = cnt[t];
        for (const char* s = cb[t]; s < cb[t + 1] && o < want;) {
            const char* nl = (const char*)std::memchr(s, '\n', cb[t + 1] - s);
            const char* le = nl ? nl : cb[t + 1];
            if (!(s < le && s[0] == '%') && !blank_line(s, le)) {
                const char* q = s;
                int64_t r = 0, c = 0;
                if (!parse_i64(q, le, &r) || !parse_i64(q, le, &c)) bad |= 1;
                else if (r < 1 || r > *nrows || c < 1 || c > *ncols) bad |= 2;
                rows[o] = r - 1;
                cols[o] = c - 1;
                ++o;
            }
            s = nl ? nl + 1 : cb[t + 1];
        }
    }
    if (cnt[nt] < want) return fail(GBK_INVALID_VALUE, "Error: unexpected end of graph file! Aborting...");
    if (bad & 1) return fail(GBK_INVALID_VALUE, "malformed .mtx entry line");
    if (bad & 2) return fail(GBK_INVALID_INDEX, ".mtx entry outside the declared %lld x %lld matrix",
                             (long long)*nrows, (long long)*ncols);
    return GBK_SUCCESS;
}

int read_binary(const char* path, int64_t* nrows, int64_t* ncols, HostCoo& out) {
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return fail(GBK_INVALID_VALUE, "Error: Can not open .mtx file (%s)", path);
    int64_t h[3];
    if (std::fread(h, sizeof(int64_t), 3, fp) != 3) {
        std::fclose(fp);
        return fail(GBK_INVALID_VALUE, "Error! Unexpected end of binary file");
    }
    if (h[0] < 0 || h[1] < 0 || h[2] < 0) {
        std::fclose(fp);
        return fail(GBK_INVALID_VALUE, "malformed .mtxbin header");
    }
    *nrows = h[0];
    *ncols = h[1];
    if (out.alloc(h[2]) != GBK_SUCCESS) {
        std::fclose(fp);
        return GBK_OUT_OF_MEMORY;
    }
    // chunks of (row, col) pairs, split into the two arrays in parallel
    const int64_t chunk = 1 << 22;
    int64_t* buf = (int64_t*)std::malloc(sizeof(int64_t) * 2 * (size_t)std::min<int64_t>(std::max<int64_t>(h[2], 1), chunk));
    if (!buf) {
        std::fclose(fp);
        return fail(GBK_OUT_OF_MEMORY, "host allocation failed");
    }
    int bad = 0;
    bool short_file = false;
    for (int64_t e0 = 0; e0 < h[2]; e0 += chunk) {
        const int64_t k = std::min(chunk, h[2] - e0);
        if (std::fread(buf, sizeof(int64_t), 2 * (size_t)k, fp) != 2 * (size_t)k) {
            short_file = true;
            break;
        }
#pragma omp parallel for schedule(static) reduction(| : bad)
        for (int64_t e = 0; e < k; ++e) {
            const int64_t r = buf[2 * e], c = buf[2 * e + 1];
            if (r < 1 || r > h[0] || c < 1 || c > h[1]) bad = 1;
            out.rows[e0 + e] = r - 1;
            out.cols[e0 + e] = c - 1;
        }
    }
    std::free(buf);
    std::fclose(fp);
    if (short_file) return fail(GBK_INVALID_VALUE, "Error! Unexpected end of binary file");
    if (bad)
        return fail(GBK_INVALID_INDEX, ".mtxbin entry outside the declared %lld x %lld matrix", (long long)h[0],
                    (long long)h[1]);
    return GBK_SUCCESS;
}

}  // namespace

namespace gbk {

// Reads either format (by suffix): 0-based entries in file order, in arrays
// allocated with malloc that the caller owns (*rows, *cols; free()).
int mtx_read(const char* path, int64_t* nrows, int64_t* ncols, int64_t* nnz, int64_t** rows, int64_t** cols) {
    if (!path) return fail(GBK_NULL_POINTER, "null file name");
    const std::string name(path);
    HostCoo h;
    int rc;
    if (ends_with(name, "mtxbin")) rc = read_binary(path, nrows, ncols, h);
    else if (ends_with(name, "mtx")) rc = read_text(path, nrows, ncols, h);
    else return fail(GBK_INVALID_VALUE, "Unsupported matrix file format. can be either .mtx or .mtx.bin");
    if (rc != GBK_SUCCESS) return rc;
    *nnz = h.nnz;
    *rows = h.rows;
    *cols = h.cols;
    h.rows = h.cols = nullptr;
    return GBK_SUCCESS;
}

}  // namespace gbk

extern "C" int gbk_mtx_read(const char* path, int64_t* nrows, int64_t* ncols, int64_t* nnz, int64_t** rows,
                            int64_t** cols) {
    if (!nrows || !ncols || !nnz || !rows || !cols) return fail(GBK_NULL_POINTER, "null argument");
    return gbk::mtx_read(path, nrows, ncols, nnz, rows, cols);
}

extern "C" int gbk_mtx_write(const char* path, int64_t nrows, int64_t ncols, const int64_t* rows,
                             const int64_t* cols, int64_t nnz) {
    if (!path || (nnz > 0 && (!rows || !cols))) return fail(GBK_NULL_POINTER, "null argument");
    if (nrows < 0 || ncols < 0 || nnz < 0) return fail(GBK_INVALID_VALUE, "negative size");
    for (int64_t e = 0; e < nnz; ++e)
        if (rows[e] < 0 || rows[e] >= nrows || cols[e] < 0 || cols[e] >= ncols)
            return fail(GBK_INVALID_INDEX, "entry %lld outside the %lld x %lld matrix", (long long)e,
                        (long long)nrows, (long long)ncols);
    const std::string name(path);
    const bool bin = ends_with(name, "mtxbin");
    if (!bin && !ends_with(name, "mtx"))
        return fail(GBK_INVALID_VALUE, "Unsupported matrix file format. can be either .mtx or .mtx.bin");
    FILE* fp = std::fopen(path, bin ? "wb" : "w");
    if (!fp) return fail(GBK_INVALID_VALUE, "cannot create %s: %s", path, std::strerror(errno));
    bool ok = true;
    if (bin) {
        const int64_t h[3] = {nrows, ncols, nnz};
        ok = std::fwrite(h, sizeof(int64_t), 3, fp) == 3;
        std::vector<int64_t> buf;
        const int64_t chunk = 1 << 20;
        for (int64_t e0 = 0; ok && e0 < nnz; e0 += chunk) {
            const int64_t e1 = std::min(nnz, e0 + chunk);
            buf.resize(2 * (e1 - e0));
            for (int64_t e = e0; e < e1; ++e) {
                buf[2 * (e - e0)] = rows[e] + 1;
                buf[2 * (e - e0) + 1] = cols[e] + 1;
            }
            ok = std::fwrite(buf.data(), sizeof(int64_t), buf.size(), fp) == buf.size();
        }
    } else {
        ok = std::fprintf(fp, "%%%%MatrixMarket matrix coordinate pattern general\n%lld %lld %lld\n",
                          (long long)nrows, (long long)ncols, (long long)nnz) > 0;
        // chunks formatted in parallel, written in order
        const int64_t chunk = 1 << 20;
        const int nt = std::max(1, std::min(omp_get_max_threads(), 64));
        std::vector<std::vector<char>> bufs(nt);
        for (int64_t g0 = 0; ok && g0 < nnz; g0 += chunk * nt) {
#pragma omp parallel num_threads(nt)
            {
                const int t = omp_get_thread_num();
                std::vector<char>& out = bufs[t];
                out.clear();
                const int64_t e0 = g0 + (int64_t)t * chunk, e1 = std::min(nnz, e0 + chunk);
                char tmp[48];
                for (int64_t e = e0; e < e1; ++e) {
                    int len = 0;
                    len += fmt_u64(tmp + len, (uint64_t)(rows[e] + 1));
                    tmp[len++] = ' ';
                    len += fmt_u64(tmp + len, (uint64_t)(cols[e] + 1));
                    tmp[len++] = '\n';
                    out.insert(out.end(), tmp, tmp + len);
                }
            }
            for (int t = 0; ok && t < nt; ++t)
                if (!bufs[t].empty()) ok = std::fwrite(bufs[t].data(), 1, bufs[t].size(), fp) == bufs[t].size();
        }
    }
    if (std::fclose(fp) != 0) ok = false;
    if (!ok) return fail(GBK_PANIC, "writing %s failed", path);
    return GBK_SUCCESS;
}

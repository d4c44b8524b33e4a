// MatrixMarket ingest (extension API, include/spike_pstrsv.h).
//
// Same semantics as the reference's matrixMarketToCSR (src/misc.c:109-245):
// coordinate files with real / integer / double values, 1-based indices;
// symm 'Y' expands a symmetric file by mirroring every off-diagonal entry,
// 'S' / 'N' read the entries as given; entries are sorted by (row, column);
// uplo 'L' keeps row >= column, 'U' keeps row <= column, 'N' keeps all; the
// result is zero-based CSR with n = the file's column count.  Differences,
// all on inputs the reference mishandles: every off-diagonal entry of a
// symmetric file is mirrored (the reference assumes exactly n diagonal
// entries, misc.c:143), rows without entries keep their (empty) place in
// ia, and malformed files are reported instead of read past.
#include "spike_pstrsv.h"

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string m_err;

int mfail(int code, const std::string &m)
{
    m_err = m;
    return code;
}

struct Entry {
    int r, c;
    double v;
};

}  // namespace

extern "C" {

int spike_read_matrix_market(const char *path, char uplo, char symm, int *n_out, double **a_out, int **ia_out,
                             int **ja_out)
{
    if (!path || !n_out || !a_out || !ia_out || !ja_out) return mfail(SPIKE_ERR_ARG, "NULL argument");
    *a_out = nullptr;
    *ia_out = *ja_out = nullptr;
    *n_out = 0;
    uplo = (char)std::toupper((unsigned char)uplo);
    symm = (char)std::toupper((unsigned char)symm);
    if (uplo != 'L' && uplo != 'U' && uplo != 'N') return mfail(SPIKE_ERR_ARG, "uplo must be L, U or N");
    if (symm != 'Y' && symm != 'S' && symm != 'N' && symm != 'A')
        return mfail(SPIKE_ERR_ARG, "symm must be Y, S, N or A");
    FILE *fp = std::fopen(path, "r");
    if (!fp) return mfail(SPIKE_ERR_ARG, std::string("cannot open ") + path);
    char line[4096];
    bool sym_banner = false, have_size = false;
    long m = 0, ncol = 0, nnz = 0;
    while (std::fgets(line, sizeof line, fp)) {
        if (line[0] == '%') {
            std::string s(line);
            for (char &ch : s) ch = (char)std::tolower((unsigned char)ch);
            if (s.rfind("%%matrixmarket", 0) == 0) {
                if (s.find("coordinate") == std::string::npos || s.find("complex") != std::string::npos ||
                    s.find("pattern") != std::string::npos) {
                    std::fclose(fp);
                    return mfail(SPIKE_ERR_UNSUPPORTED, "only real/integer coordinate MatrixMarket files are read");
                }
                sym_banner = s.find("symmetric") != std::string::npos;
            }
            continue;
        }
        if (std::sscanf(line, "%ld %ld %ld", &m, &ncol, &nnz) == 3) {
            have_size = true;
            break;
        }
        bool blank = true;
        for (const char *p = line; *p; p++)
            if (!std::isspace((unsigned char)*p)) blank = false;
        if (!blank) break;
    }
    if (!have_size || m < 1 || ncol < 1 || nnz < 0 || ncol > 0x7ffffffe || m > 0x7ffffffe) {
        std::fclose(fp);
        return mfail(SPIKE_ERR_MATRIX, "missing or invalid size line");
    }
    if (symm == 'A') symm = sym_banner ? 'Y' : 'N';
    std::vector<Entry> e;
    e.reserve((size_t)nnz * (symm == 'Y' ? 2 : 1));
    for (long k = 0; k < nnz; k++) {
        long r, c;
        double v;
        if (std::fscanf(fp, "%ld %ld %lf", &r, &c, &v) != 3) {
            std::fclose(fp);
            return mfail(SPIKE_ERR_MATRIX, "entry " + std::to_string(k + 1) + " is missing or malformed");
        }
        if (r < 1 || r > m || c < 1 || c > ncol) {
            std::fclose(fp);
            return mfail(SPIKE_ERR_MATRIX, "entry " + std::to_string(k + 1) + " is out of range");
        }
        e.push_back(Entry{(int)(r - 1), (int)(c - 1), v});
        if (symm == 'Y' && r != c) e.push_back(Entry{(int)(c - 1), (int)(r - 1), v});
    }
    std::fclose(fp);
    std::stable_sort(e.begin(), e.end(),
                     [](const Entry &x, const Entry &y) { return x.r != y.r ? x.r < y.r : x.c < y.c; });
    const int n = (int)ncol;
    std::vector<Entry> kept;
    kept.reserve(e.size());
    for (const Entry &x : e) {
        if (x.r >= n) continue;   // rows beyond the column count have no place in an n x n CSR
        if (uplo == 'L' && x.r < x.c) continue;
        if (uplo == 'U' && x.r > x.c) continue;
        kept.push_back(x);
    }
    if (kept.size() > 0x7fffffff) return mfail(SPIKE_ERR_UNSUPPORTED, "more than 2^31-1 entries");
    int *ia = (int *)std::malloc(sizeof(int) * ((size_t)n + 1));
    int *ja = (int *)std::malloc(sizeof(int) * std::max<size_t>(kept.size(), 1));
    double *a = (double *)std::malloc(sizeof(double) * std::max<size_t>(kept.size(), 1));
    if (!ia || !ja || !a) {
        std::free(ia);
        std::free(ja);
        std::free(a);
        return mfail(SPIKE_ERR_NOMEM, "host allocation failed");
    }
    std::memset(ia, 0, sizeof(int) * ((size_t)n + 1));
    for (size_t k = 0; k < kept.size(); k++) {
        ia[kept[k].r + 1]++;
        ja[k] = kept[k].c;
        a[k] = kept[k].v;
    }
    for (int i = 0; i < n; i++) ia[i + 1] += ia[i];
    *n_out = n;
    *a_out = a;
    *ia_out = ia;
    *ja_out = ja;
    return SPIKE_OK;
}

void spike_free(void *p) { std::free(p); }

const char *spike_mm_last_error(void) { return m_err.c_str(); }

}  // extern "C"

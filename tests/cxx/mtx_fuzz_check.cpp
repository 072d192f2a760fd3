// Host-side robustness check of the Matrix-Market reader (csrc/mtx.cpp), no
// GPU.  Every fixture under the given directory is read with each flag
// setting: the bad_* files must be refused with an error status, the others
// read.  Then each readable fixture is mutated many times (truncated at a
// random byte, random bytes overwritten, a random line repeated or dropped)
// and read again: a read either fails with an error status or returns
// entries inside the declared shape whose rows dspmm_csr_rowptr accepts
// (sorted, in range).  Built with AddressSanitizer by
// tests/test_mtx_fuzz.py, so an out-of-bounds access or a leak on any of
// these inputs fails the test.
#include <dirent.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "dspmm.h"

static uint64_t g_s = 1;
static uint64_t rnd() {
    g_s = g_s * 6364136223846793005ull + 1442695040888963407ull;
    return g_s >> 33;
}

static bool slurp(const std::string &path, std::string &out) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    char buf[4096];
    size_t n;
    out.clear();
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) out.append(buf, n);
    std::fclose(f);
    return true;
}

static bool spit(const std::string &path, const std::string &s) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    std::fwrite(s.data(), 1, s.size(), f);
    std::fclose(f);
    return true;
}

// 0: refused, 1: read and consistent, -1: read but inconsistent
static int read_check(const std::string &path, unsigned flags) {
    dspmm_coo c{};
    if (dspmm_mtx_read(path.c_str(), flags, &c) != DSPMM_STATUS_SUCCESS) return 0;
    int ok = c.nr >= 0 && c.nc >= 0 && c.nnz >= 0 && (c.nnz == 0 || (c.rowIdx && c.colIdx && c.values));
    for (int i = 0; ok && i < c.nnz; i++)
        if (c.rowIdx[i] < 0 || c.rowIdx[i] >= c.nr || c.colIdx[i] < 0 || c.colIdx[i] >= c.nc) ok = 0;
    if (ok) {
        std::vector<int> rp((size_t)c.nr + 1);
        ok = dspmm_csr_rowptr(c.nr, c.nnz, c.rowIdx, rp.data()) == DSPMM_STATUS_SUCCESS &&
             rp[(size_t)c.nr] == c.nnz;
    }
    dspmm_coo_free(&c);
    return ok ? 1 : -1;
}

static std::string mutate(const std::string &s) {
    std::string m = s;
    switch (rnd() % 4) {
    case 0:  // truncate
        m.resize((size_t)(rnd() % (s.size() + 1)));
        break;
    case 1: {  // overwrite a few bytes
        const char pool[] = "0123456789 \n-+.eE%xZ";
        const int k = 1 + (int)(rnd() % 4);
        for (int i = 0; i < k && !m.empty(); i++) m[(size_t)(rnd() % m.size())] = pool[rnd() % (sizeof pool - 1)];
        break;
    }
    default: {  // repeat or drop a line
        std::vector<size_t> starts{0};
        for (size_t i = 0; i + 1 < m.size(); i++)
            if (m[i] == '\n') starts.push_back(i + 1);
        const size_t k = (size_t)(rnd() % starts.size());
        const size_t b = starts[k], e = k + 1 < starts.size() ? starts[k + 1] : m.size();
        const std::string line = m.substr(b, e - b);
        if (rnd() % 2) m.insert(e, line);
        else m.erase(b, e - b);
        break;
    }
    }
    return m;
}

int main(int argc, char **argv) {
    if (argc < 3) {
        std::printf("usage: %s FIXTURE_DIR TMP_DIR\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1], tmp = std::string(argv[2]) + "/fuzz.mtx";
    std::vector<std::string> files;
    if (DIR *d = opendir(dir.c_str())) {
        while (dirent *e = readdir(d)) {
            const std::string n = e->d_name;
            if (n.size() > 4 && n.compare(n.size() - 4, 4, ".mtx") == 0) files.push_back(n);
        }
        closedir(d);
    }
    int bad = 0, good = 0, fuzzed = 0, refused = 0;
    for (const auto &n : files) {
        const bool expect_bad = n.rfind("bad_", 0) == 0;
        for (unsigned flags : {0u, DSPMM_MTX_KEEP_DIAGONAL, DSPMM_MTX_ALIGN_VALUES,
                               DSPMM_MTX_KEEP_DIAGONAL | DSPMM_MTX_ALIGN_VALUES}) {
            const int r = read_check(dir + "/" + n, flags);
            if (r < 0 || (r == 1) == expect_bad) {
                std::printf("%s flags %u: %s\n", n.c_str(), flags, r < 0 ? "inconsistent" : "wrong outcome");
                return 1;
            }
        }
        (expect_bad ? bad : good)++;
        if (expect_bad) continue;
        std::string text;
        if (!slurp(dir + "/" + n, text)) return 1;
        g_s = 1000 + text.size();
        for (int k = 0; k < 150; k++) {
            if (!spit(tmp, mutate(text))) return 1;
            const int r = read_check(tmp, (unsigned)(k % 4));
            if (r < 0) {
                std::printf("%s mutation %d: read but inconsistent\n", n.c_str(), k);
                spit(tmp + ".kept", mutate(text));
                return 1;
            }
            refused += r == 0;
            fuzzed++;
        }
    }
    std::remove(tmp.c_str());
    std::printf("fixtures: %d read, %d refused as expected; %d mutations, %d refused, none inconsistent\n",
                good, bad, fuzzed, refused);
    return good > 0 && bad > 0 ? 0 : 1;
}

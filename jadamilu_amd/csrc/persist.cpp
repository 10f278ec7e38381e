// persist.cpp -- on-disk formats: the multilevel factor and MatrixMarket.
//
// * A persisted factor holds everything the apply reads (the reference's
//   linked list of DAMGlevelmat levels, src/include/ilupack.h:197-211:
//   permutation, scaling, strict triangular factor, coupling block, pivots,
//   dense final level), so a host factorisation can be computed once and
//   reused by later runs (jdhip_ildl_load + jdhip_ildl_attach, then a solve
//   with ICNTL(2) = -1, the reference's re-entry mode for a kept
//   preconditioner, PJD.F:34-70).
//   Layout (little-endian): "JDHIPFAC", u32 version (1), u32 flags (bit 0
//   FP32 values, bit 1 |D| conversion, bit 2 L_CB coupling), i64 nnz(A_upper),
//   i32 levels, i32 dense order m, f64 fill, i32 negative pivots, f64
//   condest; per level i32 n, i32 nB, i64 nnz(L_B), i64 nnz(coupling),
//   perm[n] i32, scal[n], L_B (ptr[nB+1] i32, idx i32, val), coupling
//   (ptr[n-nB+1], idx, val), d[nB], doff[nB] (2x2 blocks), dpair[nB] i32;
//   then the dense inverse m*m f64; values
//   (scal, val, d) in the factor's precision; last a u64 FNV-1a checksum of
//   every byte before it.
// * MatrixMarket (coordinate, real, symmetric or general-with-symmetric
//   content) to and from the reference's input format: 1-based upper
//   triangle CSR with every diagonal entry stored (PJD.F:34-70).
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/jadamilu_amd.h"
#include "ildl_handle.hpp"

namespace {

constexpr char kMagic[8] = {'J', 'D', 'H', 'I', 'P', 'F', 'A', 'C'};
constexpr uint32_t kVersion = 1;

struct Fnv {
    uint64_t h = 1469598103934665603ULL;
    void add(const void *p, size_t n) {
        const unsigned char *b = static_cast<const unsigned char *>(p);
        for (size_t i = 0; i < n; ++i) {
            h ^= b[i];
            h *= 1099511628211ULL;
        }
    }
};

struct Writer {
    std::FILE *f;
    Fnv sum;
    bool ok = true;
    void raw(const void *p, size_t n) {
        if (!ok || n == 0) return;
        sum.add(p, n);
        ok = std::fwrite(p, 1, n, f) == n;
    }
    template <typename X>
    void put(X v) { raw(&v, sizeof(X)); }
    template <typename X>
    void vec(const std::vector<X> &v) { raw(v.data(), sizeof(X) * v.size()); }
};

struct Reader {
    std::FILE *f;
    Fnv sum;
    bool ok = true;
    void raw(void *p, size_t n) {
        if (!ok || n == 0) return;
        ok = std::fread(p, 1, n, f) == n;
        if (ok) sum.add(p, n);
    }
    template <typename X>
    X get() {
        X v{};
        raw(&v, sizeof(X));
        return v;
    }
    template <typename X>
    void vec(std::vector<X> &v, long n) {
        if (n < 0 || n > (1L << 40)) {
            ok = false;
            return;
        }
        v.resize((size_t)n);
        raw(v.data(), sizeof(X) * (size_t)n);
    }
};

template <typename T>
void write_ml(Writer &w, const jdhip::IldlMultilevel<T> &ml) {
    for (const auto &L : ml.levels) {
        w.put<int32_t>(L.n);
        w.put<int32_t>(L.nB);
        w.put<int64_t>((int64_t)L.lb_idx.size());
        w.put<int64_t>((int64_t)L.lc_idx.size());
        w.vec(L.perm);
        w.vec(L.scal);
        w.vec(L.lb_ptr);
        w.vec(L.lb_idx);
        w.vec(L.lb_val);
        w.vec(L.lc_ptr);
        w.vec(L.lc_idx);
        w.vec(L.lc_val);
        w.vec(L.d);
        w.vec(L.doff);
        w.vec(L.dpair);
    }
    w.vec(ml.dinv);
}

template <typename T>
bool read_ml(Reader &r, jdhip::IldlMultilevel<T> &ml, int nlev, int m) {
    ml.levels.resize(nlev);
    for (auto &L : ml.levels) {
        L.n = r.get<int32_t>();
        L.nB = r.get<int32_t>();
        const int64_t nb = r.get<int64_t>(), nc = r.get<int64_t>();
        if (!r.ok || L.n < 0 || L.nB < 0 || L.nB > L.n || nb < 0 || nc < 0) return false;
        const int nC = L.n - L.nB;
        r.vec(L.perm, L.n);
        r.vec(L.scal, L.n);
        r.vec(L.lb_ptr, L.nB + 1);
        r.vec(L.lb_idx, nb);
        r.vec(L.lb_val, nb);
        r.vec(L.lc_ptr, nC + 1);
        r.vec(L.lc_idx, nc);
        r.vec(L.lc_val, nc);
        r.vec(L.d, L.nB);
        r.vec(L.doff, L.nB);
        r.vec(L.dpair, L.nB);
        if (!r.ok) return false;
        for (int b = 0; b < L.nB; ++b) {
            const int q = L.dpair[b];
            if (q != -1 && (q < 0 || q >= L.nB || std::abs(q - b) != 1 || L.dpair[q] != b))
                return false;
        }
        // structural checks: the apply indexes with these arrays
        if (L.lb_ptr[0] != 0 || L.lb_ptr[L.nB] != nb || L.lc_ptr[0] != 0 || L.lc_ptr[nC] != nc)
            return false;
        for (int i = 0; i < L.nB; ++i)
            if (L.lb_ptr[i + 1] < L.lb_ptr[i]) return false;
        for (int c = 0; c < nC; ++c)
            if (L.lc_ptr[c + 1] < L.lc_ptr[c]) return false;
        for (int v : L.lb_idx)
            if (v < 0 || v >= L.nB) return false;
        for (int v : L.lc_idx)
            if (v < 0 || v >= L.nB) return false;
        std::vector<char> seen(L.n, 0);
        for (int v : L.perm) {
            if (v < 0 || v >= L.n || seen[v]) return false;
            seen[v] = 1;
        }
    }
    ml.m = m;
    r.vec(ml.dinv, (long)m * m);
    // level sizes chain: each level's input is the previous level's C part
    for (size_t l = 1; l < ml.levels.size(); ++l)
        if (ml.levels[l].n != ml.levels[l - 1].n - ml.levels[l - 1].nB) return false;
    if (!ml.levels.empty() && m > 0 && m != ml.levels.back().n - ml.levels.back().nB) return false;
    return r.ok;
}

}  // namespace

extern "C" {

int jdhip_ildl_save(jdhip_ildl *f, const char *path) {
    if (!f || !path) return -1;
    std::FILE *fp = std::fopen(path, "wb");
    if (!fp) return -2;
    Writer w{fp};
    const jdhip::IldlHost &h = f->h;
    const int nlev = h.single ? (int)h.ff.levels.size() : (int)h.fd.levels.size();
    const int m = h.single ? h.ff.m : h.fd.m;
    const int coupling = h.single ? h.ff.coupling : h.fd.coupling;
    w.raw(kMagic, sizeof(kMagic));
    w.put<uint32_t>(kVersion);
    w.put<uint32_t>((h.single ? 1u : 0u) | (h.spd ? 2u : 0u) | (coupling ? 4u : 0u));
    w.put<int64_t>(h.nnz_upper);
    w.put<int32_t>(nlev);
    w.put<int32_t>(m);
    w.put<double>(h.res.fill);
    w.put<int32_t>(h.res.nneg);
    w.put<double>(h.res.condest);
    if (h.single) write_ml(w, h.ff);
    else write_ml(w, h.fd);
    const uint64_t s = w.sum.h;
    w.ok = w.ok && std::fwrite(&s, sizeof(s), 1, fp) == 1;
    const bool ok = w.ok;
    if (std::fclose(fp) != 0 || !ok) return -2;
    return 0;
}

jdhip_ildl *jdhip_ildl_load(const char *path, int *ierr) {
    int dummy = 0;
    if (!ierr) ierr = &dummy;
    *ierr = -1;
    if (!path) return nullptr;
    std::FILE *fp = std::fopen(path, "rb");
    if (!fp) {
        *ierr = -2;
        return nullptr;
    }
    Reader r{fp};
    char magic[8];
    r.raw(magic, sizeof(magic));
    const uint32_t version = r.get<uint32_t>();
    const uint32_t flags = r.get<uint32_t>();
    auto *f = new jdhip_ildl();
    jdhip::IldlHost &h = f->h;
    bool ok = r.ok && std::memcmp(magic, kMagic, 8) == 0 && version == kVersion;
    if (ok) {
        h.single = flags & 1u;
        h.spd = (flags & 2u) != 0;
        h.nnz_upper = r.get<int64_t>();
        const int nlev = r.get<int32_t>();
        const int m = r.get<int32_t>();
        h.res.fill = r.get<double>();
        h.res.nneg = r.get<int32_t>();
        h.res.condest = r.get<double>();
        ok = r.ok && nlev >= 0 && nlev < 4096 && m >= 0 && m < 100000;
        if (ok) {
            ok = h.single ? read_ml(r, h.ff, nlev, m) : read_ml(r, h.fd, nlev, m);
            (h.single ? h.ff.coupling : h.fd.coupling) = (flags & 4u) ? 1 : 0;
            h.res.nlev = nlev + (m > 0 ? 1 : 0);
        }
    }
    if (ok) {
        const uint64_t expect = r.sum.h;
        uint64_t got = 0;
        ok = std::fread(&got, sizeof(got), 1, fp) == 1 && got == expect;
    }
    std::fclose(fp);
    if (!ok) {
        delete f;
        *ierr = -3;
        return nullptr;
    }
    *ierr = 0;
    return f;
}

// ---- MatrixMarket ----------------------------------------------------------

static int mm_parse(const char *path, int *n_out, std::vector<long> *ri, std::vector<long> *ci,
                    std::vector<double> *v) {
    std::ifstream in(path);
    if (!in) return -2;
    std::string line;
    if (!std::getline(in, line)) return -3;
    std::string lower = line;
    for (auto &c : lower) c = (char)std::tolower((unsigned char)c);
    if (lower.rfind("%%matrixmarket", 0) != 0 || lower.find("matrix") == std::string::npos ||
        lower.find("coordinate") == std::string::npos)
        return -3;
    const bool pattern = lower.find("pattern") != std::string::npos;
    const bool symmetric = lower.find("symmetric") != std::string::npos;
    const bool general = lower.find("general") != std::string::npos;
    if (lower.find("complex") != std::string::npos || lower.find("hermitian") != std::string::npos ||
        lower.find("skew") != std::string::npos || !(symmetric || general))
        return -4;
    long rows = -1, cols = -1, nnz = -1;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '%') continue;
        std::istringstream ss(line);
        ss >> rows >> cols >> nnz;
        break;
    }
    if (rows <= 0 || rows != cols || nnz < 0 || rows > 2147483647L) return -3;
    *n_out = (int)rows;
    ri->clear();
    ci->clear();
    v->clear();
    for (long k = 0; k < nnz; ++k) {
        long i, j;
        double x = 1.0;
        if (!(in >> i >> j)) return -3;
        if (!pattern && !(in >> x)) return -3;
        if (i < 1 || j < 1 || i > rows || j > rows) return -3;
        if (general && i > j) continue;  // the upper triangle of a symmetric matrix
        if (symmetric && i < j) return -3;  // symmetric files store the lower triangle
        ri->push_back(std::min(i, j));
        ci->push_back(std::max(i, j));
        v->push_back(x);
    }
    return 0;
}

// rows of the upper triangle, duplicates summed, diagonal present in every row
static void mm_to_upper(int n, const std::vector<long> &ri, const std::vector<long> &ci,
                        const std::vector<double> &v, std::vector<int> &ia, std::vector<int> &ja,
                        std::vector<double> &a) {
    std::vector<std::vector<std::pair<int, double>>> rows(n);
    for (int i = 0; i < n; ++i) rows[i].emplace_back(i + 1, 0.0);
    for (size_t k = 0; k < ri.size(); ++k) rows[ri[k] - 1].emplace_back((int)ci[k], v[k]);
    ia.assign(n + 1, 1);
    ja.clear();
    a.clear();
    for (int i = 0; i < n; ++i) {
        auto &r = rows[i];
        std::stable_sort(r.begin(), r.end(),
                         [](const auto &x, const auto &y) { return x.first < y.first; });
        for (size_t t = 0; t < r.size(); ++t) {
            if (t > 0 && r[t].first == r[t - 1].first) {
                a.back() += r[t].second;
                continue;
            }
            ja.push_back(r[t].first);
            a.push_back(r[t].second);
        }
        ia[i + 1] = (int)ja.size() + 1;
    }
}

int jdhip_mm_info(const char *path, int *n, long *nnz_upper) {
    if (!path || !n || !nnz_upper) return -1;
    std::vector<long> ri, ci;
    std::vector<double> v;
    const int rc = mm_parse(path, n, &ri, &ci, &v);
    if (rc) return rc;
    std::vector<int> ia, ja;
    std::vector<double> a;
    mm_to_upper(*n, ri, ci, v, ia, ja, a);
    *nnz_upper = (long)ja.size();
    return 0;
}

int jdhip_mm_read(const char *path, int n, long nnz_cap, int *ia, int *ja, double *a) {
    if (!path || !ia || !ja || !a) return -1;
    std::vector<long> ri, ci;
    std::vector<double> v;
    int nf = 0;
    const int rc = mm_parse(path, &nf, &ri, &ci, &v);
    if (rc) return rc;
    if (nf != n) return -5;
    std::vector<int> vi, vj;
    std::vector<double> va;
    mm_to_upper(n, ri, ci, v, vi, vj, va);
    if ((long)vj.size() > nnz_cap) return -5;
    std::copy(vi.begin(), vi.end(), ia);
    std::copy(vj.begin(), vj.end(), ja);
    std::copy(va.begin(), va.end(), a);
    return 0;
}

int jdhip_mm_write(const char *path, int n, const int *ia, const int *ja, const double *a) {
    if (!path || n < 0 || !ia || !ja || !a) return -1;
    std::FILE *fp = std::fopen(path, "w");
    if (!fp) return -2;
    const long nnz = (long)ia[n] - 1;
    bool ok = std::fprintf(fp, "%%%%MatrixMarket matrix coordinate real symmetric\n") > 0 &&
              std::fprintf(fp, "%d %d %ld\n", n, n, nnz) > 0;
    for (int i = 0; ok && i < n; ++i)
        for (int k = ia[i] - 1; ok && k < ia[i + 1] - 1; ++k) {
            const int r = i + 1, c = ja[k];
            // lower triangle as the symmetric format stores it; 17 digits
            // round-trip every double
            ok = std::fprintf(fp, "%d %d %.17g\n", std::max(r, c), std::min(r, c), a[k]) > 0;
        }
    if (std::fclose(fp) != 0) ok = false;
    return ok ? 0 : -2;
}

}  // extern "C"

// Seeded synthetic problem generators and the MatrixMarket reader.
//
// The five generators produce the five workloads this library is measured on
// (sizes and distributions stated next to each); every random draw goes
// through Rng in a fixed order, so a (generator, seed) pair gives the same
// matrix on every machine.
#include "problems.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace spai {

// ---------------------------------------------------------------- Rng ----
static uint64_t splitmix64(uint64_t &x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

Rng::Rng(uint64_t seed) {
    uint64_t x = seed;
    for (auto &w : s_) w = splitmix64(x);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

uint64_t Rng::next() {
    uint64_t result = rotl(s_[1] * 5, 7) * 9;
    uint64_t t = s_[1] << 17;
    s_[2] ^= s_[0];
    s_[3] ^= s_[1];
    s_[1] ^= s_[2];
    s_[0] ^= s_[3];
    s_[2] ^= t;
    s_[3] = rotl(s_[3], 45);
    return result;
}

double Rng::uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }

double Rng::normal() {
    if (has_spare_) {
        has_spare_ = false;
        return spare_;
    }
    double u1 = uniform(), u2 = uniform();
    if (u1 <= 0.0) u1 = 1e-300;
    double r = std::sqrt(-2.0 * std::log(u1));
    double th = 6.283185307179586 * u2;
    spare_ = r * std::sin(th);
    has_spare_ = true;
    return r * std::cos(th);
}

int Rng::below(int n) { return (int)((next() >> 33) % (uint64_t)n); }

// ------------------------------------------------------------ helpers ----
namespace {
// Row-wise builder: rows appended in order, (col, val) pairs per row.
struct RowBuilder {
    int n;
    std::vector<long long> rowptr{0};
    std::vector<int> col;
    std::vector<double> val;
    explicit RowBuilder(int n_) : n(n_) { rowptr.reserve((size_t)n + 1); }
    void push(int c, double v) {
        col.push_back(c);
        val.push_back(v);
    }
    void end_row() { rowptr.push_back((long long)col.size()); }
    HostCSC to_csc() const {
        HostCSC A;
        A.n = n;
        size_t nnz = col.size();
        A.colptr.assign((size_t)n + 1, 0);
        for (size_t p = 0; p < nnz; ++p) A.colptr[(size_t)col[p] + 1]++;
        for (int c = 0; c < n; ++c) A.colptr[(size_t)c + 1] += A.colptr[(size_t)c];
        std::vector<int> cur(A.colptr.begin(), A.colptr.end() - 1);
        A.rowidx.resize(nnz);
        A.val.resize(nnz);
        for (int r = 0; r < n; ++r) {
            for (long long p = rowptr[(size_t)r]; p < rowptr[(size_t)r + 1]; ++p) {
                int q = cur[(size_t)col[(size_t)p]]++;
                A.rowidx[(size_t)q] = r;
                A.val[(size_t)q] = val[(size_t)p];
            }
        }
        return A;
    }
};

std::vector<double> row_sums(const HostCSC &A) {
    std::vector<double> ones((size_t)A.n, 1.0), y((size_t)A.n);
    csc_matvec(A, ones.data(), y.data());
    return y;
}

std::string fmt(const char *f, double a = 0, double b = 0, double c = 0, double d = 0) {
    char buf[512];
    std::snprintf(buf, sizeof buf, f, a, b, c, d);
    return buf;
}
}  // namespace

HostCSC csc_from_rows(int n, const std::vector<std::vector<std::pair<int, double>>> &rows) {
    RowBuilder rb(n);
    for (int r = 0; r < n; ++r) {
        for (auto &e : rows[(size_t)r]) rb.push(e.first, e.second);
        rb.end_row();
    }
    return rb.to_csc();
}

void csc_matvec(const HostCSC &A, const double *x, double *y) {
    for (int i = 0; i < A.n; ++i) y[i] = 0.0;
    for (int c = 0; c < A.n; ++c)
        for (int p = A.colptr[(size_t)c]; p < A.colptr[(size_t)c + 1]; ++p)
            y[A.rowidx[(size_t)p]] += A.val[(size_t)p] * x[c];
}

// -------------------------------------------------------------- grids ----
// -eps*Laplace(u) + v.grad(u) on the unit square, m x m interior points,
// h = 1/(m+1), eps = 1/Pe, 5-point diffusion, first-order upwind convection,
// Dirichlet boundary eliminated, every row scaled by h^2.  Unknown (i, j) is
// row i + m*j, i along x.
HostCSC convection_diffusion_2d(int m, double peclet, double vx, double vy) {
    int n = m * m;
    double h = 1.0 / (m + 1);
    double eps = 1.0 / peclet;
    RowBuilder rb(n);
    double cw = -eps - (vx > 0 ? h * vx : 0.0), ce = -eps + (vx < 0 ? h * vx : 0.0);
    double cs = -eps - (vy > 0 ? h * vy : 0.0), cn = -eps + (vy < 0 ? h * vy : 0.0);
    double cc = 4.0 * eps + h * (std::fabs(vx) + std::fabs(vy));
    for (int j = 0; j < m; ++j) {
        for (int i = 0; i < m; ++i) {
            int r = i + m * j;
            if (j > 0) rb.push(r - m, cs);
            if (i > 0) rb.push(r - 1, cw);
            rb.push(r, cc);
            if (i < m - 1) rb.push(r + 1, ce);
            if (j < m - 1) rb.push(r + m, cn);
            rb.end_row();
        }
    }
    return rb.to_csc();
}

// Cell-centred finite volumes for -div(kappa grad u) on an m^3 grid, interior
// faces weighted by the harmonic mean of the two cells, boundary faces
// (Dirichlet, half-cell distance) by 2*kappa of the cell.  Cell (i,j,k) is row
// i + m*(j + m*k).
HostCSC diffusion_3d(int m, const std::vector<double> &kappa) {
    int n = m * m * m;
    RowBuilder rb(n);
    auto K = [&](int i, int j, int k) { return kappa[(size_t)i + (size_t)m * ((size_t)j + (size_t)m * k)]; };
    auto hm = [](double a, double b) { return 2.0 * a * b / (a + b); };
    for (int k = 0; k < m; ++k)
        for (int j = 0; j < m; ++j)
            for (int i = 0; i < m; ++i) {
                int r = i + m * (j + m * k);
                double kc = K(i, j, k), diag = 0.0;
                double f[6];
                f[0] = k > 0 ? hm(kc, K(i, j, k - 1)) : 0.0;
                f[1] = j > 0 ? hm(kc, K(i, j - 1, k)) : 0.0;
                f[2] = i > 0 ? hm(kc, K(i - 1, j, k)) : 0.0;
                f[3] = i < m - 1 ? hm(kc, K(i + 1, j, k)) : 0.0;
                f[4] = j < m - 1 ? hm(kc, K(i, j + 1, k)) : 0.0;
                f[5] = k < m - 1 ? hm(kc, K(i, j, k + 1)) : 0.0;
                diag += (k > 0 ? f[0] : 2.0 * kc) + (j > 0 ? f[1] : 2.0 * kc) + (i > 0 ? f[2] : 2.0 * kc);
                diag += (i < m - 1 ? f[3] : 2.0 * kc) + (j < m - 1 ? f[4] : 2.0 * kc) +
                        (k < m - 1 ? f[5] : 2.0 * kc);
                if (k > 0) rb.push(r - m * m, -f[0]);
                if (j > 0) rb.push(r - m, -f[1]);
                if (i > 0) rb.push(r - 1, -f[2]);
                rb.push(r, diag);
                if (i < m - 1) rb.push(r + 1, -f[3]);
                if (j < m - 1) rb.push(r + m, -f[4]);
                if (k < m - 1) rb.push(r + m * m, -f[5]);
                rb.end_row();
            }
    return rb.to_csc();
}

// --------------------------------------------------------- generators ----
// n = 4096; per row 4 off-diagonal columns drawn uniformly (self-loops and
// repeats redrawn), values U(-1,1), diagonal 1.5 x the row's absolute
// off-diagonal sum; b = A * ones.
Problem generate_random_diag_dominant(uint64_t seed) {
    const int n = 4096, k = 4;
    Rng rng(seed);
    RowBuilder rb(n);
    for (int i = 0; i < n; ++i) {
        int cols[k];
        double vals[k];
        int got = 0;
        while (got < k) {
            int c = rng.below(n);
            bool dup = (c == i);
            for (int t = 0; t < got; ++t) dup = dup || cols[t] == c;
            if (!dup) cols[got++] = c;
        }
        double s = 0.0;
        for (int t = 0; t < k; ++t) {
            vals[t] = 2.0 * rng.uniform() - 1.0;
            s += std::fabs(vals[t]);
        }
        // emit the row with its entries in ascending column order
        std::vector<std::pair<int, double>> row;
        for (int t = 0; t < k; ++t) row.push_back({cols[t], vals[t]});
        row.push_back({i, 1.5 * s});
        std::sort(row.begin(), row.end());
        for (auto &e : row) rb.push(e.first, e.second);
        rb.end_row();
    }
    Problem P;
    P.name = "random4096";
    P.seed = seed;
    P.systems.push_back(rb.to_csc());
    P.rhs.push_back(row_sums(P.systems[0]));
    P.target.emplace_back();
    P.notes = "n=4096, 4 off-diagonals/row U(-1,1), a_ii=1.5*sum|a_ij|, b=A*1";
    return P;
}

// 512 x 512 conv-diff, v = (1, 0.5), Pe = 10 (eps = 0.1), b ~ U(0,1).
Problem generate_convection_diffusion_512(uint64_t seed) {
    Rng rng(seed);
    Problem P;
    P.name = "convdiff512";
    P.seed = seed;
    P.systems.push_back(convection_diffusion_2d(512, 10.0, 1.0, 0.5));
    std::vector<double> b((size_t)P.systems[0].n);
    for (auto &v : b) v = rng.uniform();
    P.rhs.push_back(std::move(b));
    P.target.emplace_back();
    P.notes = "m=512, h=1/513, eps=1/Pe=0.1, v=(1,0.5), upwind, scaled by h^2, b~U(0,1)";
    return P;
}

// Power-law rows: off-diagonal count per row = min(Zipf(s=2), 1000) (inverse
// CDF over 1..999, the remaining mass on 1000), distinct uniform columns != i,
// values N(0,1), diagonal 1.2 x the absolute row sum, b ~ U(0,1).
Problem generate_power_law(uint64_t seed, int n) {
    Rng rng(seed);
    const int cap = 1000;
    std::vector<double> cdf(cap);
    const double zeta2 = 1.6449340668482264;
    double acc = 0.0;
    for (int k = 1; k < cap; ++k) {
        acc += 1.0 / ((double)k * k * zeta2);
        cdf[(size_t)k - 1] = acc;
    }
    cdf[(size_t)cap - 1] = 2.0;  // everything above 999 is capped at 1000
    RowBuilder rb(n);
    std::vector<int> cols;
    std::vector<std::pair<int, double>> row;
    for (int i = 0; i < n; ++i) {
        double u = rng.uniform();
        int d = (int)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin()) + 1;
        if (d > cap) d = cap;
        if (d > n - 1) d = n - 1;
        cols.clear();
        while ((int)cols.size() < d) {
            int need = d - (int)cols.size();
            for (int t = 0; t < need; ++t) {
                int c = rng.below(n);
                if (c != i) cols.push_back(c);
            }
            std::sort(cols.begin(), cols.end());
            cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
        }
        row.clear();
        double s = 0.0;
        for (int c : cols) {
            double v = rng.normal();
            s += std::fabs(v);
            row.push_back({c, v});
        }
        row.push_back({i, 1.2 * s});
        std::sort(row.begin(), row.end());
        for (auto &e : row) rb.push(e.first, e.second);
        rb.end_row();
    }
    Problem P;
    P.name = "powerlaw" + std::to_string(n);
    P.seed = seed;
    P.systems.push_back(rb.to_csc());
    std::vector<double> b((size_t)n);
    for (auto &v : b) v = rng.uniform();
    P.rhs.push_back(std::move(b));
    P.target.emplace_back();
    P.notes = "off-diag/row=min(Zipf(2),1000), N(0,1) values, a_ii=1.2*sum|a_ij|, b~U(0,1)";
    return P;
}

// nsys independent systems, sizes N_s = round(exp(U(ln 2e3, ln 2e5))).  Even
// s: 2D conv-diff on round(sqrt(N_s))^2 with Pe ~ U(1,20), v = (1, 0.5).  Odd
// s: 3D diffusion on round(cbrt(N_s))^3 with cell kappa = exp(N(0,1)).
// b_s ~ U(0,1).
Problem generate_batched(uint64_t seed, int nsys) {
    Rng rng(seed);
    Problem P;
    P.name = "batched" + std::to_string(nsys);
    P.seed = seed;
    const double lo = std::log(2e3), hi = std::log(2e5);
    for (int s = 0; s < nsys; ++s) {
        double Ns = std::exp(lo + (hi - lo) * rng.uniform());
        HostCSC A;
        if (s % 2 == 0) {
            int m = std::max(2, (int)std::lround(std::sqrt(Ns)));
            double pe = 1.0 + 19.0 * rng.uniform();
            A = convection_diffusion_2d(m, pe, 1.0, 0.5);
        } else {
            int m = std::max(2, (int)std::lround(std::cbrt(Ns)));
            std::vector<double> kappa((size_t)m * m * m);
            for (auto &k : kappa) k = std::exp(rng.normal());
            A = diffusion_3d(m, kappa);
        }
        std::vector<double> b((size_t)A.n);
        for (auto &v : b) v = rng.uniform();
        P.systems.push_back(std::move(A));
        P.rhs.push_back(std::move(b));
        P.target.emplace_back();
    }
    P.notes = "N_s=round(exp(U(ln2e3,ln2e5))); even: 2D conv-diff Pe~U(1,20) v=(1,0.5); "
              "odd: 3D FV diffusion kappa=exp(N(0,1)) harmonic faces; b~U(0,1)";
    return P;
}

// 64^3 diffusion with kappa = exp(g), g a Gaussian random field: white noise
// smoothed by a periodic separable Gaussian filter of width sigma = l/sqrt(2)
// (field correlation exp(-r^2/(2 l^2)), l = 8 cells), then shifted and
// scaled to zero mean and unit variance.  b, x_t ~ N(0,1).
Problem generate_diffusion_grf_64(uint64_t seed, int m) {
    Rng rng(seed);
    const size_t N = (size_t)m * m * m;
    std::vector<double> f(N), g(N);
    for (auto &v : f) v = rng.normal();
    const double ell = 8.0, sigma = ell / std::sqrt(2.0);
    std::vector<double> w((size_t)m);
    for (int d = 0; d < m; ++d) {
        int dd = std::min(d, m - d);
        w[(size_t)d] = std::exp(-(double)dd * dd / (2.0 * sigma * sigma));
    }
    // three periodic 1D convolutions (x, then y, then z)
    for (int axis = 0; axis < 3; ++axis) {
        size_t stride = axis == 0 ? 1 : (axis == 1 ? (size_t)m : (size_t)m * m);
        for (size_t idx = 0; idx < N; ++idx) {
            size_t coord = (idx / stride) % (size_t)m;
            size_t base = idx - coord * stride;
            double s = 0.0;
            for (int t = 0; t < m; ++t) {
                size_t src = base + (size_t)t * stride;
                int d = (int)((coord + (size_t)m - (size_t)t) % (size_t)m);
                s += w[(size_t)d] * f[src];
            }
            g[idx] = s;
        }
        std::swap(f, g);
    }
    double mean = 0.0, var = 0.0;
    for (double v : f) mean += v;
    mean /= (double)N;
    for (double v : f) var += (v - mean) * (v - mean);
    double sd = std::sqrt(var / (double)N);
    std::vector<double> kappa(N);
    for (size_t i = 0; i < N; ++i) kappa[i] = std::exp((f[i] - mean) / sd);
    Problem P;
    P.name = "grf" + std::to_string(m);
    P.seed = seed;
    P.systems.push_back(diffusion_3d(m, kappa));
    std::vector<double> b(N), xt(N);
    for (auto &v : b) v = rng.normal();
    for (auto &v : xt) v = rng.normal();
    P.rhs.push_back(std::move(b));
    P.target.push_back(std::move(xt));
    P.notes = fmt("m=%g, kappa=exp(GRF), filter sigma=%.4f cells (corr length 8), b,x_t~N(0,1)", m, sigma);
    return P;
}

Problem generate_config(int config, uint64_t seed) {
    switch (config) {
    case 1: return generate_random_diag_dominant(seed);
    case 2: return generate_convection_diffusion_512(seed);
    case 3: return generate_power_law(seed);
    case 4: return generate_batched(seed);
    case 5: return generate_diffusion_grf_64(seed);
    default: throw std::invalid_argument("config must be 1..5");
    }
}

// ----------------------------------------------------- MatrixMarket ----
HostCSC read_matrix_market(const std::string &path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open " + path);
    std::string line;
    if (!std::getline(in, line)) throw std::runtime_error("empty file " + path);
    bool symmetric = line.find("symmetric") != std::string::npos;
    // skip comments up to the size line
    do {
        if (!std::getline(in, line)) throw std::runtime_error("no size line in " + path);
    } while (!line.empty() && line[0] == '%');
    long long nr = 0, nc = 0, nz = 0;
    {
        std::istringstream ss(line);
        if (!(ss >> nr >> nc >> nz)) throw std::runtime_error("bad size line in " + path);
    }
    if (nr <= 0 || nc <= 0 || nr > (1LL << 31) - 2 || nc > (1LL << 31) - 2 || nz < 0)
        throw std::runtime_error("bad matrix dimensions in " + path);
    std::vector<std::vector<int>> ci((size_t)nc);
    std::vector<std::vector<double>> cv((size_t)nc);
    int ntok = -1;
    long long read = 0;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '%') continue;
        std::istringstream ss(line);
        std::vector<std::string> tok;
        std::string t;
        while (ss >> t) tok.push_back(t);
        if (tok.empty()) continue;
        if (ntok < 0) ntok = (int)tok.size();
        if ((int)tok.size() < 2) throw std::runtime_error("bad entry line in " + path);
        long long i = std::stoll(tok[0]), j = std::stoll(tok[1]);
        if (i < 1 || i > nr || j < 1 || j > nc) throw std::runtime_error("index out of range in " + path);
        double v = 1.0;
        bool valued = ntok >= 3;
        if (valued) {
            if ((int)tok.size() < 3) throw std::runtime_error("missing value in " + path);
            v = std::stod(tok[2]);
        }
        ++read;
        if (valued && std::fabs(v) == 0.0) {
            if (!symmetric && read >= nz) break;
            continue;
        }
        ci[(size_t)j - 1].push_back((int)(i - 1));
        cv[(size_t)j - 1].push_back(v);
        if (symmetric && i != j) {
            ci[(size_t)i - 1].push_back((int)(j - 1));
            cv[(size_t)i - 1].push_back(v);
        }
        if (!symmetric && read >= nz) break;
    }
    if (!symmetric && read < nz) throw std::runtime_error("truncated entries in " + path);
    HostCSC A;
    A.n = (int)nc;
    A.colptr.assign((size_t)nc + 1, 0);
    for (long long c = 0; c < nc; ++c) A.colptr[(size_t)c + 1] = A.colptr[(size_t)c] + (int)ci[(size_t)c].size();
    A.rowidx.reserve((size_t)A.colptr[(size_t)nc]);
    A.val.reserve((size_t)A.colptr[(size_t)nc]);
    for (long long c = 0; c < nc; ++c) {
        A.rowidx.insert(A.rowidx.end(), ci[(size_t)c].begin(), ci[(size_t)c].end());
        A.val.insert(A.val.end(), cv[(size_t)c].begin(), cv[(size_t)c].end());
    }
    if (nr != nc) throw std::runtime_error("matrix is not square: " + path);
    return A;
}

}  // namespace spai

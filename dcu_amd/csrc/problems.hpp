// Host-side problem containers, seeded synthetic generators and the
// MatrixMarket reader.  Pure C++ (no HIP): the generators feed both the GPU
// path and the tests' CPU oracle with identical inputs.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace spai {

// Host CSC with the same meaning as CSC_Matrix (0-based, square).
struct HostCSC {
    int n = 0;
    std::vector<int> colptr;   // n + 1
    std::vector<int> rowidx;   // nnz
    std::vector<double> val;   // nnz
    long long nnz() const { return (long long)rowidx.size(); }
};

// One or more independent systems with right-hand sides.
struct Problem {
    std::string name;
    uint64_t seed = 0;
    std::vector<HostCSC> systems;     // 1 system except the batched config
    std::vector<std::vector<double>> rhs;     // per system
    std::vector<std::vector<double>> target;  // per system, empty unless the loss needs one
    std::string notes;                // generator parameters actually used
};

// xoshiro256** seeded through splitmix64; uniform doubles in [0,1) use the top
// 53 bits, normals use Box-Muller with the second value cached.
class Rng {
public:
    explicit Rng(uint64_t seed);
    uint64_t next();
    double uniform();
    double normal();
    int below(int n);  // uniform integer in [0, n)
private:
    uint64_t s_[4];
    bool has_spare_ = false;
    double spare_ = 0.0;
};

// Build a CSC matrix from per-row (column, value) lists; columns in a row
// must be distinct.  Rows of the CSC are ordered ascending inside each column.
HostCSC csc_from_rows(int n, const std::vector<std::vector<std::pair<int, double>>> &rows);

// y = A x with A in CSC.
void csc_matvec(const HostCSC &A, const double *x, double *y);

Problem generate_random_diag_dominant(uint64_t seed);        // 4096, 4 off-diagonals/row
Problem generate_convection_diffusion_512(uint64_t seed);     // 512x512 upwind conv-diff
Problem generate_power_law(uint64_t seed, int n = 1000000);   // Zipf(2)-capped rows
Problem generate_batched(uint64_t seed, int nsys = 64);       // 2D conv-diff / 3D diffusion mix
Problem generate_diffusion_grf_64(uint64_t seed, int m = 64); // 3D lognormal-GRF diffusion

// Configuration index 1..5 in the order above.
Problem generate_config(int config, uint64_t seed);

// MatrixMarket coordinate file with the reader semantics of the reference
// (include/common/read.h:1296-1321): explicit zeros dropped for valued files,
// `symmetric` headers mirrored entry by entry in file order, file order kept
// inside each column, pattern files read as 1.0, the first of two values used.
// Throws std::runtime_error on malformed input.
HostCSC read_matrix_market(const std::string &path);

// Individual grids, exposed for tests.
HostCSC convection_diffusion_2d(int m, double peclet, double vx, double vy);
HostCSC diffusion_3d(int m, const std::vector<double> &kappa);

}  // namespace spai

// cc_lattice.h -- host-side lattice bookkeeping of the CC engine: seeded configurations, reciprocal
// basis, cell volume, the GetKs wavevector set, and the split of a lattice constraint set into
// "segments" (runs of consecutive last-axis indices) that the GPU kernels walk with a phase recurrence.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace ccgpu {

constexpr double kPi = 3.1415926535897932384626433832795028841971693993751058209;  // core/etc.h:9
constexpr int kMaxDim = 4;                                                            // core/etc.h:12

// 32-bit Mersenne twister with the standard seeding (boost::mt19937 / std::mt19937, core/RandomGenerator.h:16).
class MT19937 {
public:
    explicit MT19937(uint32_t seed) { Seed(seed); }
    void Seed(uint32_t seed);
    uint32_t Next();
    // RandomDouble(): (u - min) / ((double)(max - min) + 1)  (core/RandomGenerator.cpp:24-27)
    double NextDouble() { return (double)Next() / ((double)4294967295u + 1); }

private:
    uint32_t mt_[624];
    int idx_;
};

// Insert's wrap into [0,1): subtract the floor twice (core/PeriodicCellList.h:527-530).
double WrapUnit(double x);

// N random particles in relative coordinates, d draws per particle in order (apps/CCOptimization/main.cpp:207-215).
void RandomRelativeConfiguration(uint32_t seed, int64_t N, int d, double *s_out);

// Reciprocal basis b_i with b_i . a_j = 2 pi delta_ij, computed with the Gram-Schmidt steps of
// GetReciprocalBasisVector (core/PeriodicCellList.h:999-1031).  a, b: d rows of d doubles.
void ReciprocalBasis(int d, const double *a, double *b);

// |det a| (PeriodicVolume, core/GeometryVector.cpp:47-77, LU with partial pivoting).
double CellVolume(int d, const double *a);

// getK(chi, rho, d) (apps/CCOptimization/main.cpp:2018-2021).
double GetK(double chi, double rho, int d);

// GetKs(config, K, K, 1) (core/PeriodicCellList.cpp:509-735): the half-space reciprocal lattice vectors with
// |k| < K (d >= 2; d == 1 keeps k <= K), sorted by (|k|^2, k_x).  Fills Cartesian k (M x d) and integer
// indices n (M x d).
void GetKs(int d, const double *a, double K, std::vector<double> &k, std::vector<int32_t> &n);

// Integer indices of Cartesian wavevectors k (M x d) on the reciprocal lattice of basis a:
// n_j = round(k . a_j / 2 pi).  Returns false if any k is farther than a rounding tolerance from
// sum_j n_j b_j (the constraint set is then not a lattice set and the general kernels are used).
bool LatticeIndices(int d, const double *a, int64_t M, const double *k, std::vector<int32_t> &n);

// Segments: constraints grouped by the first d-1 indices, each group cut into runs of consecutive
// last-axis index of at most S entries.  Slot q = seg*S + t holds constraint slot_k[q] (or -1 for padding)
// with index (head[0..d-2], head[d-1] + t).  Segments are ordered longest first, so the lanes of one
// wavefront walk runs of similar length.
struct SegmentPlan {
    int S = 0;
    int64_t nseg = 0;
    std::vector<int32_t> head;     // nseg x 4 (start mode, zero padded)
    std::vector<int32_t> len;      // nseg
    std::vector<int32_t> slot_k;   // nseg*S
};
SegmentPlan BuildSegments(int d, int64_t M, const int32_t *n, int S);

}  // namespace ccgpu

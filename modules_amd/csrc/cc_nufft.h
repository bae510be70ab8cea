// cc_nufft.h -- device layout of the NUFFT backend of the lattice path (the reference's FINUFFT rho backend,
// potential/CollectiveCoordinatePotential.h:479-751, 1053-1132, with its Type 2 forces).
//
// A lattice constraint k = sum_c n_c b_c has k.r = n.x with x = 2 pi s (s relative coordinates), so
//   rho(n) = sum_j exp(i n.x_j)                         (Type 1, strengths 1)
//   G_c(j) = Im sum_k n_kc (2 V_k / vol) rho_k exp(-i n_k.x_j)   (Type 2; F_j = -sum_c G_c(j) b_c)
// are non-uniform FFTs over the mode box |n_c| <= K_c.  Method (written from the published algorithm of Barnett,
// Magland and af Klinteberg, "A parallel non-uniform fast Fourier transform library based on an 'exponential of
// semicircle' kernel", SIAM J. Sci. Comput. 41 (2019)): spread each particle with the kernel
// phi(z) = exp(beta (sqrt(1 - z^2) - 1)) over w points per axis of a fine grid of nf_c >= 2 (2 K_c + 1) points,
// FFT, and divide mode n by p(n) = prod_c (w/2) int_{-1}^{1} phi(z) cos(pi w n_c z / nf_c) dz; Type 2 runs the
// same steps backwards.  w = ceil(log10(1/eps)) + 1 and beta = 2.30 w (upsampling 2).
//
// The force pass needs two real sums per particle and axis pair; with the constraints' conjugates the mode
// coefficients f_c(n) = n_c c_n / 2 (c_{-n} = conj c_n) make sum_n f_c(n) exp(-i n.x) = i G_c, so one complex
// transform of f_0 + i f_1 gives G_0 = Im u and G_1 = -Re u (a second transform carries f_2 in 3D).
//
// Determinism: particles are binned by output tile (counting sort, each bin sorted by particle index); each tile
// spreads its own particles, in index order, into a padded tile (the tile plus the kernel's reach), and each grid
// point then sums the padded tiles that cover it in a fixed tile order -- every sum has a fixed order.
#pragma once
#include <cstdint>
#include <hip/hip_runtime.h>

namespace ccgpu {

constexpr int kNufftMaxW = 16;   // kernel width cap (eps down to ~1e-15)
constexpr int kNufftMaxPad = 17 * 21 * 21;  // largest padded tile in LDS (3D tiles 4 x 8 x 8 with w = 13)

struct NufftDesc {
    double2 *grid;            // fine grid(s): npts complex (Type 1), then the Type 2 coefficient grid(s) (x 2 in 3D)
    double *kv;               // N x 3 x kNufftMaxW kernel values (axis c, support point o)
    int32_t *start;           // N x 4 first support point per axis (may lie outside [0, nf))
    int32_t *bin_of;          // N: output-tile bin of each particle
    int32_t *bin_start;       // nbin + 1 starts, then nbin scatter cursors
    int32_t *bin_items;       // N particle indices, by bin, index order within a bin
    const double *p;          // deconvolution factors: axis c at p + poff[c], mode m at index m + K[c]
    double *sub;              // ntile padded tiles of psize points: each tile's own particles spread locally
    int32_t nf[3];            // fine grid points per axis (axis 0 slowest; 1 beyond the dimension)
    int32_t K[3];             // mode half-width per axis
    int32_t tile[3];          // output tile (one block) per axis
    int32_t nbins[3];         // nf / tile per axis
    int32_t reach[3];         // neighbouring bins whose particles can touch a tile, per side
    int32_t poff[3];
    int32_t pad[3];           // padded tile extent per axis: tile + w (1 beyond the dimension)
    int32_t w, nbin, npts, ntile;
    int32_t hlo, psize;       // padding below a tile (floor(w/2)); points of one padded tile
    double beta;
};

}  // namespace ccgpu

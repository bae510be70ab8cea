/*
 * cc_oracle.c -- CPU oracle for the collective-coordinate (CC) stealthy-potential path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker, not the product: only tests/, the smoke
 * entry and the CPU-baseline leg of bench.py load it.  The product (modules_amd/) never calls it.
 *
 * It restates, in plain C with the reference's operation order, the FP64 semantics of:
 *   - RandomGenerator (boost::mt19937 + RandomDouble)      core/RandomGenerator.cpp:24-27
 *   - random configuration GetUnitCubicBox/Rescale/Insert   apps/CCOptimization/main.cpp:207-215,
 *                                                           core/PeriodicCellList.h:88-98, 517-549
 *   - GetReciprocalBasisVector                              core/PeriodicCellList.h:999-1031
 *   - GetKs (d>=2 lattice walk, half-space rule, sort; d=1) core/PeriodicCellList.cpp:509-735
 *   - getK                                                  apps/CCOptimization/main.cpp:2018-2021
 *   - ShiftedCCPotential::GetRho_Direct / Energy / Force    potential/CollectiveCoordinatePotential.h:1255-1394
 *   - ShiftedCCPotential_S2::Energy / Force (target S(k))   potential/CollectiveCoordinatePotential.h:1686-1741
 *   - ShiftedCCPotential_withWeight::GetRho (weighted rho)  potential/CollectiveCoordinatePotential.h:1520-1550
 *   - ShiftedCCPotential_S2_MultiConfig rho / E / F        potential/CollectiveCoordinatePotential.h:1777-1879
 *   - RepulsiveCCPotential soft core energy / forces         apps/CCOptimization/RepulsiveCCPotential.h:61-145
 *     (brute force over periodic images instead of the cell-list walk; same pair terms)
 *   - MultiStealthyPotential energy / forces (species CC      apps/CCOptimization/MultiStealthyPotential.cpp:682-833
 *     terms over index lists, radius contact term)
 *   - StructureOptimization Objective / FillParticles / GetParam and the 3-pass drivers
 *                                                           potential/StructureOptimization.cpp:339-350, 447-578,
 *                                                           710-776, 881-979, 1046-1084
 * The L-BFGS itself is injected as a function pointer with LuksanPLIS's signature
 * (potential/LuksanPLIS.h:35-36); tests pass the reference's own LuksanPLIS compiled under oracle/_ref.
 *
 * Build: see oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off, no -march=native, as config.mk:128-143).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define ORC_PI 3.1415926535897932384626433832795028841971693993751058209 /* core/etc.h:9 */
#define ORC_MAXDIM 4                                                      /* core/etc.h:12 */

/* ------------------------------------------------------------------------------------------------
 * MT19937 (the standardized 32-bit Mersenne twister; boost::mt19937 and std::mt19937 share the
 * seeding routine and the output sequence).  RandomGenerator::seed(int) seeds with the int value.
 * ---------------------------------------------------------------------------------------------- */
typedef struct { uint32_t mt[624]; int idx; } orc_mt;

void orc_mt_seed(orc_mt *g, uint32_t s)
{
	g->mt[0] = s;
	for (int i = 1; i < 624; i++)
		g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
	g->idx = 624;
}

uint32_t orc_mt_next(orc_mt *g)
{
	if (g->idx >= 624) {
		for (int i = 0; i < 624; i++) {
			uint32_t y = (g->mt[i] & 0x80000000u) | (g->mt[(i + 1) % 624] & 0x7fffffffu);
			uint32_t v = g->mt[(i + 397) % 624] ^ (y >> 1);
			if (y & 1u) v ^= 0x9908b0dfu;
			g->mt[i] = v;
		}
		g->idx = 0;
	}
	uint32_t y = g->mt[g->idx++];
	y ^= y >> 11;
	y ^= (y << 7) & 0x9d2c5680u;
	y ^= (y << 15) & 0xefc60000u;
	y ^= y >> 18;
	return y;
}

/* RandomDouble = (gen()-min)/((double)(max-min)+1) with min=0, max=2^32-1 (core/RandomGenerator.cpp:24-27) */
double orc_mt_double(orc_mt *g)
{
	return (double)(orc_mt_next(g) - 0u) / ((double)(4294967295u - 0u) + 1);
}

/* First `count` raw outputs after seeding (for known-answer tests). */
void orc_mt_sequence(uint32_t seed, int64_t count, uint32_t *out)
{
	orc_mt g;
	orc_mt_seed(&g, seed);
	for (int64_t i = 0; i < count; i++) out[i] = orc_mt_next(&g);
}

/* Insert's wrap: subtract the floor twice (core/PeriodicCellList.h:527-530). */
static double orc_wrap(double x)
{
	double y = x - floor(x);
	y -= floor(y);
	return y;
}

/* Random configuration: N particles, each drawing d RandomDouble()s in order, wrapped as Insert does
 * (apps/CCOptimization/main.cpp:207-215 with RandomGenerator(seed)).  Output: relative coords N x d. */
void orc_make_config(uint32_t seed, int64_t N, int d, double *s_out)
{
	orc_mt g;
	orc_mt_seed(&g, seed);
	for (int64_t i = 0; i < N; i++)
		for (int j = 0; j < d; j++)
			s_out[i * d + j] = orc_wrap(orc_mt_double(&g));
}

/* Continue an existing generator: used for the MultiRun re-randomisation (main.cpp:1863-1869). */
void orc_make_config_from(orc_mt *g, int64_t N, int d, double *s_out)
{
	for (int64_t i = 0; i < N; i++)
		for (int j = 0; j < d; j++)
			s_out[i * d + j] = orc_wrap(orc_mt_double(g));
}

/* ------------------------------------------------------------------------------------------------
 * GeometryVector helpers over MaxDimension=4 components, zero padded (core/GeometryVector.h:192-206).
 * ---------------------------------------------------------------------------------------------- */
static double gv_dot(const double *a, const double *b)
{
	double r = 0;
	for (int i = 0; i < ORC_MAXDIM; i++) r += a[i] * b[i];
	return r;
}

/* Reciprocal basis vector i (core/PeriodicCellList.h:999-1031): Gram-Schmidt of the other basis
 * vectors, projection of a_i onto their orthogonal complement, then scale by 2*pi/(result . a_i).
 * a: d rows of d doubles.  b_out: d rows of d doubles. */
void orc_reciprocal_basis(int d, const double *a, double *b_out)
{
	double A[ORC_MAXDIM][ORC_MAXDIM];
	memset(A, 0, sizeof(A));
	for (int i = 0; i < d; i++)
		for (int j = 0; j < d; j++) A[i][j] = a[i * d + j];
	for (int i = 0; i < d; i++) {
		double res[ORC_MAXDIM], ortho[ORC_MAXDIM][ORC_MAXDIM];
		int no = 0;
		memcpy(res, A[i], sizeof(res));
		for (int j = 0; j < d; j++) {
			if (j == i) continue;
			double t[ORC_MAXDIM];
			memcpy(t, A[j], sizeof(t));
			for (int o = 0; o < no; o++) {
				double f = gv_dot(t, ortho[o]) / gv_dot(ortho[o], ortho[o]);
				for (int c = 0; c < ORC_MAXDIM; c++) t[c] -= ortho[o][c] * f;
			}
			memcpy(ortho[no++], t, sizeof(t));
		}
		for (int o = 0; o < no; o++) {
			double f = gv_dot(ortho[o], res) / gv_dot(ortho[o], ortho[o]);
			for (int c = 0; c < ORC_MAXDIM; c++) res[c] -= ortho[o][c] * f;
		}
		double product = gv_dot(res, A[i]);
		double scale = 2 * ORC_PI / product;
		for (int c = 0; c < ORC_MAXDIM; c++) res[c] *= scale;
		for (int j = 0; j < d; j++) b_out[i * d + j] = res[j];
	}
}

/* |det(a)| by LU with partial pivoting (core/GeometryVector.cpp:47-77 via GSL). */
double orc_volume(int d, const double *a)
{
	double m[ORC_MAXDIM][ORC_MAXDIM];
	for (int i = 0; i < d; i++)
		for (int j = 0; j < d; j++) m[i][j] = a[i * d + j];
	double det = 1.0;
	for (int c = 0; c < d; c++) {
		int p = c;
		double best = fabs(m[c][c]);
		for (int r = c + 1; r < d; r++)
			if (fabs(m[r][c]) > best) { best = fabs(m[r][c]); p = r; }
		if (p != c) {
			for (int j = 0; j < d; j++) { double t = m[c][j]; m[c][j] = m[p][j]; m[p][j] = t; }
			det = -det;
		}
		if (m[c][c] != 0.0)
			for (int r = c + 1; r < d; r++) {
				double f = m[r][c] / m[c][c];
				for (int j = c + 1; j < d; j++) m[r][j] -= f * m[c][j];
			}
	}
	for (int c = 0; c < d; c++) det *= m[c][c];
	return fabs(det);
}

/* V_d(1): HyperSphere_C (core/etc.cpp:52-66) for d = 1..4. */
static double orc_ball_c(int d)
{
	switch (d) {
	case 1: return 2;
	case 2: return ORC_PI;
	case 3: return 4.1887902047863909846;
	case 4: return 4.9348022005446793094;
	default: return pow(ORC_PI, d / 2.0) / tgamma(d / 2.0 + 1);
	}
}

/* getK(chi, rho, d) = 2 pi (2 d rho chi / V_d(1))^(1/d)  (apps/CCOptimization/main.cpp:2018-2021) */
double orc_getK(double chi, double rho, int d)
{
	return 2 * ORC_PI * pow(2 * d * rho * chi / (orc_ball_c(d) * pow(1.0, (double)d)), 1. / d);
}

typedef struct { double k[ORC_MAXDIM]; double m2; int32_t n[ORC_MAXDIM]; } orc_kv;

static int orc_kv_cmp(const void *pa, const void *pb)
{
	const orc_kv *a = (const orc_kv *)pa, *b = (const orc_kv *)pb;
	if (a->m2 != b->m2) return a->m2 < b->m2 ? -1 : 1;
	if (a->k[0] != b->k[0]) return a->k[0] < b->k[0] ? -1 : 1;
	/* ties in (|k|^2, k_x) have an unspecified order in the reference (std::sort); fix one here */
	for (int i = 1; i < ORC_MAXDIM; i++)
		if (a->k[i] != b->k[i]) return a->k[i] < b->k[i] ? -1 : 1;
	return 0;
}

/* GetKs(config, K, K, 1) for a cell with basis a (d rows).  d>=2: every reciprocal lattice vector
 * k = sum_i n_i b_i with |k|^2 < K^2 whose first nonzero n_i is positive, k accumulated in the order
 * of IterateThroughNeighbors' LatticeShift (core/PeriodicCellList.h:883-897), sorted by (|k|^2, k_x)
 * (core/PeriodicCellList.cpp:543-560, 621-626).  d=1: k = i*b for i=1,2,... while k <= K (:676-690).
 * Writes at most cap vectors (k_out: cap x d, n_out: cap x d, may be NULL) and returns the count. */
int64_t orc_get_ks(int d, const double *a, double K, int64_t cap, double *k_out, int32_t *n_out)
{
	if (K <= 0.0) return 0;
	double b[ORC_MAXDIM * ORC_MAXDIM];
	orc_reciprocal_basis(d, a, b);
	if (d == 1) {
		int64_t cnt = 0;
		for (double i = 1.0;; i += 1.0) {
			double k = i * b[0];
			if (k > K) break;
			if (cnt < cap) {
				if (k_out) k_out[cnt] = k;
				if (n_out) n_out[cnt] = (int32_t)i;
			}
			cnt++;
		}
		return cnt;
	}
	int lim[ORC_MAXDIM];
	int64_t total = 1;
	for (int i = 0; i < d; i++) {
		double an = 0;
		for (int j = 0; j < d; j++) an += a[i * d + j] * a[i * d + j];
		lim[i] = (int)ceil(K * sqrt(an) / (2. * ORC_PI)) + 1;
		total *= 2 * lim[i] + 1;
	}
	double Rc2 = fabs(K) * fabs(K);
	int64_t ncap = 1024, cnt = 0;
	orc_kv *list = (orc_kv *)malloc(sizeof(orc_kv) * ncap);
	int n[ORC_MAXDIM];
	for (int i = 0; i < d; i++) n[i] = -lim[i];
	for (int64_t it = 0; it < total; it++) {
		int keep = 0;
		for (int i = 0; i < d; i++) {
			if (n[i] > 0) { keep = 1; break; }
			if (n[i] < 0) { keep = 0; break; }
		}
		if (keep) {
			double ls[ORC_MAXDIM] = {0, 0, 0, 0};
			for (int i = 0; i < d; i++)
				for (int c = 0; c < d; c++) ls[c] += (double)n[i] * b[i * d + c];
			double sh[ORC_MAXDIM];
			for (int c = 0; c < ORC_MAXDIM; c++) sh[c] = 0.0 + ls[c];
			double m2 = gv_dot(sh, sh);
			if (m2 < Rc2 && m2 > 0.0) {
				if (cnt == ncap) { ncap *= 2; list = (orc_kv *)realloc(list, sizeof(orc_kv) * ncap); }
				memcpy(list[cnt].k, sh, sizeof(sh));
				list[cnt].m2 = m2;
				for (int c = 0; c < ORC_MAXDIM; c++) list[cnt].n[c] = c < d ? n[c] : 0;
				cnt++;
			}
		}
		n[0]++;
		for (int i = 0; i + 1 < d; i++)
			if (n[i] > lim[i]) { n[i] = -lim[i]; n[i + 1]++; }
	}
	qsort(list, (size_t)cnt, sizeof(orc_kv), orc_kv_cmp);
	for (int64_t q = 0; q < cnt && q < cap; q++)
		for (int c = 0; c < d; c++) {
			if (k_out) k_out[q * d + c] = list[q].k[c];
			if (n_out) n_out[q * d + c] = list[q].n[c];
		}
	free(list);
	return cnt;
}

/* Cartesian coordinates r = sum_i a_i s_i accumulated as UpdateCartesianCoordinates (PeriodicCellList.h:88-98). */
static void orc_cart(int d, const double *a, const double *s, double *r4)
{
	for (int c = 0; c < ORC_MAXDIM; c++) r4[c] = 0.0;
	for (int i = 0; i < d; i++)
		for (int c = 0; c < d; c++) r4[c] += a[i * d + c] * s[i];
}

static void orc_pad(int d, const double *v, double *v4)
{
	for (int c = 0; c < ORC_MAXDIM; c++) v4[c] = c < d ? v[c] : 0.0;
}

/* GetRho_Direct (CollectiveCoordinatePotential.h:1255-1286): rho_k = sum_j cos(k.r_j) + i sin(k.r_j),
 * parallel over k (each k sums its particles in order, so the bits do not depend on the thread count). */
void orc_rho(int d, int64_t N, const double *s, const double *a, int64_t M, const double *k,
             double *re, double *im, int nthreads)
{
	double *r = (double *)malloc(sizeof(double) * ORC_MAXDIM * (N > 0 ? N : 1));
	for (int64_t j = 0; j < N; j++) orc_cart(d, a, s + j * d, r + j * ORC_MAXDIM);
	if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(static)
	for (int64_t i = 0; i < M; i++) {
		double k4[ORC_MAXDIM], tr = 0, ti = 0;
		orc_pad(d, k + i * d, k4);
		for (int64_t j = 0; j < N; j++) {
			tr += cos(gv_dot(k4, r + j * ORC_MAXDIM));
			ti += sin(gv_dot(k4, r + j * ORC_MAXDIM));
		}
		re[i] = tr;
		im[i] = ti;
	}
	free(r);
}

/* ShiftedCCPotential_withWeight::GetRho (CollectiveCoordinatePotential.h:1520-1550): rho_k = sum_j w_j cos(k.r_j)
 * + i w_j sin(k.r_j), parallel over k. */
void orc_rho_weighted(int d, int64_t N, const double *s, const double *a, int64_t M, const double *k,
                      const double *w, double *re, double *im, int nthreads)
{
	double *r = (double *)malloc(sizeof(double) * ORC_MAXDIM * (N > 0 ? N : 1));
	for (int64_t j = 0; j < N; j++) orc_cart(d, a, s + j * d, r + j * ORC_MAXDIM);
	if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(static)
	for (int64_t i = 0; i < M; i++) {
		double k4[ORC_MAXDIM], tr = 0, ti = 0;
		orc_pad(d, k + i * d, k4);
		for (int64_t j = 0; j < N; j++) {
			tr += cos(gv_dot(k4, r + j * ORC_MAXDIM)) * w[j];
			ti += sin(gv_dot(k4, r + j * ORC_MAXDIM)) * w[j];
		}
		re[i] = tr;
		im[i] = ti;
	}
	free(r);
}

/* Energy (CollectiveCoordinatePotential.h:1351-1372) with one thread: sum_k V_k (Re^2+Im^2), /Volume, +Shift. */
double orc_energy(int64_t M, const double *v, const double *re, const double *im, double volume, double shift)
{
	double part = 0.0;
	for (int64_t i = 0; i < M; i++) part += v[i] * (re[i] * re[i] + im[i] * im[i]);
	double result = 0;
	result += part;
	return result / (volume) + shift;
}

/* Force / AllForce (CollectiveCoordinatePotential.h:1377-1394, 1434-1448):
 * F_j = (1/V) * ( - sum_k (2 V_k (Im cos(k.r) - Re sin(k.r))) k ), parallel over particles. */
void orc_forces(int d, int64_t N, const double *s, const double *a, int64_t M, const double *k,
                const double *v, const double *re, const double *im, double volume, double *F, int nthreads)
{
	if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(static)
	for (int64_t j = 0; j < N; j++) {
		double r4[ORC_MAXDIM], res[ORC_MAXDIM] = {0, 0, 0, 0};
		orc_cart(d, a, s + j * d, r4);
		for (int64_t i = 0; i < M; i++) {
			double k4[ORC_MAXDIM];
			orc_pad(d, k + i * d, k4);
			double c = (2.0) * v[i] * (im[i] * cos(gv_dot(k4, r4)) - re[i] * sin(gv_dot(k4, r4)));
			for (int q = 0; q < ORC_MAXDIM; q++) res[q] -= c * k4[q];
		}
		double inv = 1.0 / volume;
		for (int q = 0; q < d; q++) F[j * d + q] = res[q] * inv;
	}
}

/* ShiftedCCPotential_S2::Energy (CollectiveCoordinatePotential.h:1686-1712) with one thread:
 * sum_k V_k (S_k - S0_k)^2 + Shift, S_k = (Re^2 + Im^2) / N (no 1/Volume). */
double orc_s2_energy(int64_t M, const double *v, const double *s0, const double *re, const double *im, int64_t N,
                     double shift)
{
	double result = 0.0;
	for (int64_t i = 0; i < M; i++) {
		double s = (re[i] * re[i] + im[i] * im[i]) / (double)N;
		double ds = s - s0[i];
		result += v[i] * ds * ds;
	}
	return result + shift;
}

/* ShiftedCCPotential_S2::Force (CollectiveCoordinatePotential.h:1718-1741):
 * F_j = - sum_k (4/N) V_k (S_k - S0_k) (Im cos(k.r) - Re sin(k.r)) k, parallel over particles. */
void orc_s2_forces(int d, int64_t N, const double *s, const double *a, int64_t M, const double *k, const double *v,
                   const double *s0, const double *re, const double *im, double *F, int nthreads)
{
	if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(static)
	for (int64_t j = 0; j < N; j++) {
		double r4[ORC_MAXDIM], res[ORC_MAXDIM] = {0, 0, 0, 0};
		orc_cart(d, a, s + j * d, r4);
		for (int64_t i = 0; i < M; i++) {
			double k4[ORC_MAXDIM];
			orc_pad(d, k + i * d, k4);
			double sk = (re[i] * re[i] + im[i] * im[i]) / (double)N;
			double ds = sk - s0[i];
			double c = (4.0 / (double)N) * v[i] * ds * (im[i] * cos(gv_dot(k4, r4)) - re[i] * sin(gv_dot(k4, r4)));
			for (int q = 0; q < ORC_MAXDIM; q++) res[q] -= c * k4[q];
		}
		for (int q = 0; q < d; q++) F[j * d + q] = res[q];
	}
}

/* ShiftedCCPotential_S2_MultiConfig (CollectiveCoordinatePotential.h:1756-1901): GetRho_MultiConfig (:1777-1808,
 * rho of every sub-configuration of Np = N / nsub consecutive particles, Snow = sum |rho_c|^2 / (Np nsub)),
 * Energy (:1818-1844) and Force (:1849-1879) of every particle.  re / im: nsub x M.  Returns the energy. */
double orc_s2_multiconfig(int d, int64_t N, const double *s, const double *a, int64_t M, const double *k,
                          const double *v, const double *s0, int64_t nsub, double shift, double *re, double *im,
                          double *F, int nthreads)
{
	const int64_t np = N / nsub;
	double *snow = (double *)calloc((size_t)(M > 0 ? M : 1), sizeof(double));
	for (int64_t c = 0; c < nsub; c++) {
		orc_rho(d, np, s + c * np * d, a, M, k, re + c * M, im + c * M, nthreads);
		for (int64_t i = 0; i < M; i++) snow[i] += re[c * M + i] * re[c * M + i] + im[c * M + i] * im[c * M + i];
	}
	for (int64_t i = 0; i < M; i++) snow[i] /= (double)(np * nsub);
	double result = 0.0;
	for (int64_t i = 0; i < M; i++) {
		double ds = snow[i] - s0[i];
		result += v[i] * ds * ds;
	}
	result += shift;
	if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(static)
	for (int64_t j = 0; j < N; j++) {
		const int64_t inc = j / np;
		double r4[ORC_MAXDIM], res[ORC_MAXDIM] = {0, 0, 0, 0};
		orc_cart(d, a, s + j * d, r4);
		for (int64_t i = 0; i < M; i++) {
			double k4[ORC_MAXDIM];
			orc_pad(d, k + i * d, k4);
			double ds = snow[i] - s0[i];
			double c = (4.0 / (double)np) * v[i] * ds *
			           (im[i + inc * M] * cos(gv_dot(k4, r4)) - re[i + inc * M] * sin(gv_dot(k4, r4)));
			for (int q = 0; q < ORC_MAXDIM; q++) res[q] -= c * k4[q];
		}
		for (int q = 0; q < d; q++) F[j * d + q] = res[q] * (1.0 / (double)nsub);
	}
	free(snow);
	return result;
}

/* ------------------------------------------------------------------------------------------------
 * Soft core (RepulsiveCCPotential.h:61-145): E = e0 sum_{i<j, r<sigma} (1 - r/sigma)^2 over periodic
 * images; F_i = -(2 e0/sigma) sum_{j != i, r<sigma} (1/r - 1/sigma) (r_j + image - r_i); coincident
 * points (r < 1e-15) push along +-x.  Brute force over the image box that can reach sigma.
 * ---------------------------------------------------------------------------------------------- */
static int orc_image_reach(int d, const double *a, double sigma, int *reach)
{
	/* |n_i| <= ceil(sigma*|b_i|/(2 pi)) + 1 images along a_i suffice for |shift| < sigma after wrapping */
	double b[ORC_MAXDIM * ORC_MAXDIM];
	orc_reciprocal_basis(d, a, b);
	int64_t tot = 1;
	for (int i = 0; i < d; i++) {
		double bn = 0;
		for (int c = 0; c < d; c++) bn += b[i * d + c] * b[i * d + c];
		reach[i] = (int)ceil(sigma * sqrt(bn) / (2 * ORC_PI)) + 1;
		tot *= 2 * reach[i] + 1;
	}
	return (int)tot;
}

static void orc_image_shift(int d, const double *a, const int *reach, int idx, double *sh4)
{
	int n[ORC_MAXDIM];
	for (int i = 0; i < d; i++) {
		int w = 2 * reach[i] + 1;
		n[i] = idx % w - reach[i];
		idx /= w;
	}
	for (int c = 0; c < ORC_MAXDIM; c++) sh4[c] = 0.0;
	for (int i = 0; i < d; i++)
		for (int c = 0; c < d; c++) sh4[c] += (double)n[i] * a[i * d + c];
}

double orc_softcore_energy(int d, int64_t N, const double *s, const double *a, double sigma, double e0, int nthreads)
{
	int reach[ORC_MAXDIM];
	int nimg = orc_image_reach(d, a, sigma, reach);
	double *r = (double *)malloc(sizeof(double) * ORC_MAXDIM * (N > 0 ? N : 1));
	for (int64_t j = 0; j < N; j++) orc_cart(d, a, s + j * d, r + j * ORC_MAXDIM);
	double total = 0.0;
	if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 16) reduction(+ : total)
	for (int64_t i = 0; i < N; i++) {
		double E = 0.0;
		for (int64_t j = i + 1; j < N; j++)
			for (int q = 0; q < nimg; q++) {
				double sh[ORC_MAXDIM], dv[ORC_MAXDIM];
				orc_image_shift(d, a, reach, q, sh);
				for (int c = 0; c < ORC_MAXDIM; c++) dv[c] = (r[j * ORC_MAXDIM + c] - r[i * ORC_MAXDIM + c]) + sh[c];
				double rr = sqrt(gv_dot(dv, dv)) / sigma;
				if (rr < 1.0) E += e0 * (1.0 - rr) * (1.0 - rr);
			}
		total += E;
	}
	free(r);
	return total;
}

void orc_softcore_forces(int d, int64_t N, const double *s, const double *a, double sigma, double e0, double *F, int nthreads)
{
	int reach[ORC_MAXDIM];
	int nimg = orc_image_reach(d, a, sigma, reach);
	double *r = (double *)malloc(sizeof(double) * ORC_MAXDIM * (N > 0 ? N : 1));
	for (int64_t j = 0; j < N; j++) orc_cart(d, a, s + j * d, r + j * ORC_MAXDIM);
	if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 16)
	for (int64_t i = 0; i < N; i++) {
		double f[ORC_MAXDIM] = {0, 0, 0, 0};
		for (int64_t j = 0; j < N; j++) {
			for (int q = 0; q < nimg; q++) {
				double sh[ORC_MAXDIM], dv[ORC_MAXDIM];
				orc_image_shift(d, a, reach, q, sh);
				if (j == i) continue; /* ForceLJ skips the particle itself, all images (RepulsiveCCPotential.h:99) */
				for (int c = 0; c < ORC_MAXDIM; c++) dv[c] = (r[j * ORC_MAXDIM + c] - r[i * ORC_MAXDIM + c]) + sh[c];
				double m2 = gv_dot(dv, dv);
				if (!(m2 < sigma * sigma)) continue;
				double rr = sqrt(m2);
				if (rr < 1e-15) {
					dv[0] = (i > j) ? -1.0 : 1.0;
					for (int c = 0; c < ORC_MAXDIM; c++) f[c] = f[c] - 2.0 * e0 / sigma * dv[c];
				} else {
					double c0 = 2.0 * e0 / sigma * (1.0 / rr - 1.0 / sigma);
					for (int c = 0; c < ORC_MAXDIM; c++) f[c] = f[c] - c0 * dv[c];
				}
			}
		}
		for (int c = 0; c < d; c++) F[i * d + c] = f[c];
	}
	free(r);
}

/* ------------------------------------------------------------------------------------------------
 * Multi-species potential (apps/CCOptimization/MultiStealthyPotential.cpp).
 * Contact term with per-particle radii, sigma_ij = r_i + r_j, brute force over the periodic images that can
 * lie within the cutoff: per-particle energy slots E_i = sum_{j > i, images, |r|/sigma_ij < 1} (1 - |r|/sigma_ij)^2
 * (Energy's contact lambda, :705-720), and forces e0 * sum_{j != i, images} of ContactForce's pair terms
 * (:799-833): coincident points (|r| < 1e-15) push by -2/sigma along +-x, else 2/sigma (1/|r| - 1/sigma) (-r).
 * ---------------------------------------------------------------------------------------------- */
void orc_contact_radii(int d, int64_t N, const double *s, const double *a, const double *radii, double rc, double e0,
                       double *E_slots, double *F, int nthreads)
{
	double rmax = 0.0;
	for (int64_t i = 0; i < N; i++) rmax = radii[i] > rmax ? radii[i] : rmax;
	int reach[ORC_MAXDIM];
	int nimg = orc_image_reach(d, a, (2.0 * rmax > rc ? 2.0 * rmax : rc), reach);
	double *r = (double *)malloc(sizeof(double) * ORC_MAXDIM * (N > 0 ? N : 1));
	for (int64_t j = 0; j < N; j++) orc_cart(d, a, s + j * d, r + j * ORC_MAXDIM);
	if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 16)
	for (int64_t i = 0; i < N; i++) {
		double E = 0.0, f[ORC_MAXDIM] = {0, 0, 0, 0};
		for (int64_t j = 0; j < N; j++) {
			if (j == i) continue; /* no self pair, no self image */
			const double sigma = radii[i] + radii[j];
			for (int q = 0; q < nimg; q++) {
				double sh[ORC_MAXDIM], dv[ORC_MAXDIM];
				orc_image_shift(d, a, reach, q, sh);
				for (int c = 0; c < ORC_MAXDIM; c++) dv[c] = (r[j * ORC_MAXDIM + c] - r[i * ORC_MAXDIM + c]) + sh[c];
				const double rr = sqrt(gv_dot(dv, dv));
				if (i < j) {
					const double rs = rr / sigma;
					if (rs < 1.0) E += (1.0 - rs) * (1.0 - rs);
				}
				if (rr < 1e-15) {
					dv[0] = (i > j) ? -1.0 : 1.0;
					for (int c = 0; c < ORC_MAXDIM; c++) f[c] = f[c] - 2.0 / sigma * dv[c];
				} else if (rr < sigma) {
					const double c0 = 2.0 / sigma * (1.0 / rr - 1.0 / sigma);
					for (int c = 0; c < ORC_MAXDIM; c++) f[c] = f[c] + c0 * (-1. * dv[c]);
				}
			}
		}
		E_slots[i] = E;
		for (int c = 0; c < d; c++) F[i * d + c] = f[c] * e0;
	}
	free(r);
}

/* Energy() and AllForce() of MultiStealthyPotential (:682-761, 763-797): species s has particles idx[idx_off[s] ..
 * idx_off[s+1]) (its own order) and constraints k/v[k_off[s] .. k_off[s+1]); each species is a ShiftedCCPotential
 * over its particles in the full cell (Shift 0).  E = u0 sum_s (N_s/N) E_s + e0 sum_i E_i (TotalEnergy, :750-761);
 * F_i = contact_i + (u0 N_s/N) F_s(i) (Force, :763-797).  species_E (nspecies, may be NULL) receives the E_s. */
double orc_multi_energy_forces(int d, int64_t N, const double *s, const double *a, int nspecies, const int64_t *idx_off,
                               const int64_t *idx, const int64_t *k_off, const double *k, const double *v,
                               const double *radii, double rc, double u0, double e0, int repulsion, double *F,
                               double *species_E, int nthreads)
{
	const double V = orc_volume(d, a);
	double CC = 0.0, CP = 0.0;
	double *slots = (double *)calloc((size_t)(N > 0 ? N : 1), sizeof(double));
	double *fc = (double *)calloc((size_t)(N > 0 ? N : 1) * d, sizeof(double));
	if (repulsion) orc_contact_radii(d, N, s, a, radii, rc, e0, slots, fc, nthreads);
	for (int sp = 0; sp < nspecies; sp++) {
		const int64_t n_s = idx_off[sp + 1] - idx_off[sp], M = k_off[sp + 1] - k_off[sp];
		double *sub = (double *)malloc(sizeof(double) * (size_t)(n_s > 0 ? n_s : 1) * d);
		for (int64_t q = 0; q < n_s; q++)
			for (int c = 0; c < d; c++) sub[q * d + c] = s[idx[idx_off[sp] + q] * d + c];
		double *re = (double *)malloc(sizeof(double) * (size_t)(M > 0 ? M : 1));
		double *im = (double *)malloc(sizeof(double) * (size_t)(M > 0 ? M : 1));
		double *fs = (double *)malloc(sizeof(double) * (size_t)(n_s > 0 ? n_s : 1) * d);
		orc_rho(d, n_s, sub, a, M, k + k_off[sp] * d, re, im, nthreads);
		const double E_s = orc_energy(M, v + k_off[sp], re, im, V, 0.0);
		if (species_E) species_E[sp] = E_s;
		CC += ((double)n_s / (double)N) * E_s;
		orc_forces(d, n_s, sub, a, M, k + k_off[sp] * d, v + k_off[sp], re, im, V, fs, nthreads);
		const double w = u0 * (double)n_s / (double)N;
		for (int64_t q = 0; q < n_s; q++) {
			const int64_t i = idx[idx_off[sp] + q];
			for (int c = 0; c < d; c++) F[i * d + c] = fc[i * d + c] + fs[q * d + c] * w;
		}
		free(sub);
		free(re);
		free(im);
		free(fs);
	}
	for (int64_t i = 0; i < N; i++) CP += slots[i];
	free(slots);
	free(fc);
	return u0 * CC + e0 * CP;
}

/* ------------------------------------------------------------------------------------------------
 * Objective and the relax drivers (StructureOptimization.cpp), Switch 0, Pressure 0, MinDistance 0.
 * ---------------------------------------------------------------------------------------------- */
typedef double (*orc_objective_fn)(unsigned n, const double *x, double *grad, void *data);
typedef int (*orc_plis_fn)(int n, orc_objective_fn f, void *f_data, double *x, double *minf, double stopval,
                           int maxeval, int mf, int *evaluations, int fast);

typedef struct {
	int d;
	int64_t N, M;
	const double *a;     /* basis rows, d x d */
	const double *k;     /* M x d */
	const double *v;     /* M */
	double shift, volume;
	double sigma, e0;    /* soft core: on when |e0| > 1e-10 (RepulsiveCCPotential.h:49) */
	double rescale;
	int nthreads;
	int64_t nevals;
	double *s, *re, *im, *F, *Fc;   /* work arrays */
} orc_problem;

/* E + F of the relative coordinates in P->s; forces into P->F.  Returns the energy. */
static double orc_eval(orc_problem *P, int want_forces)
{
	int soft = fabs(P->e0) > 1e-10;
	double LJ = 0.0;
	if (soft) LJ = orc_softcore_energy(P->d, P->N, P->s, P->a, P->sigma, P->e0, P->nthreads);
	orc_rho(P->d, P->N, P->s, P->a, P->M, P->k, P->re, P->im, P->nthreads);
	double CC = orc_energy(P->M, P->v, P->re, P->im, P->volume, P->shift);
	if (want_forces) {
		orc_forces(P->d, P->N, P->s, P->a, P->M, P->k, P->v, P->re, P->im, P->volume, P->F, P->nthreads);
		if (soft) {
			orc_softcore_forces(P->d, P->N, P->s, P->a, P->sigma, P->e0, P->Fc, P->nthreads);
			for (int64_t i = 0; i < P->N * P->d; i++) P->F[i] = P->F[i] + P->Fc[i];
		}
	}
	return LJ + CC;
}

/* Objective (StructureOptimization.cpp:447-578, Switch 0): FillParticles (x/rescale, wrapped as Insert),
 * energy, and grad[i*d+j] = (-1) F_i . a_j / rescale. */
static double orc_objective(unsigned n, const double *x, double *grad, void *data)
{
	orc_problem *P = (orc_problem *)data;
	(void)n;
	for (int64_t i = 0; i < P->N * P->d; i++) P->s[i] = orc_wrap(x[i] / P->rescale);
	double f = orc_eval(P, grad != NULL);
	if (grad) {
		for (int64_t i = 0; i < P->N; i++)
			for (int j = 0; j < P->d; j++) {
				double F4[ORC_MAXDIM], a4[ORC_MAXDIM];
				orc_pad(P->d, P->F + i * P->d, F4);
				orc_pad(P->d, P->a + j * P->d, a4);
				grad[i * P->d + j] = (-1.0) * gv_dot(F4, a4) / P->rescale;
			}
	}
	P->nevals++;
	return f;
}

static void orc_problem_alloc(orc_problem *P)
{
	size_t nd = (size_t)(P->N * P->d > 0 ? P->N * P->d : 1), m = (size_t)(P->M > 0 ? P->M : 1);
	P->s = (double *)malloc(sizeof(double) * nd);
	P->F = (double *)malloc(sizeof(double) * nd);
	P->Fc = (double *)malloc(sizeof(double) * nd);
	P->re = (double *)malloc(sizeof(double) * m);
	P->im = (double *)malloc(sizeof(double) * m);
}

static void orc_problem_free(orc_problem *P)
{
	free(P->s); free(P->F); free(P->Fc); free(P->re); free(P->im);
}

/* Total energy (soft core + CC) and forces of relative coordinates s (one call of the objective at rescale 1). */
double orc_total_energy_forces(int d, int64_t N, const double *s, const double *a, int64_t M, const double *k,
                               const double *v, double shift, double sigma, double e0, double *F_out, int nthreads)
{
	orc_problem P;
	memset(&P, 0, sizeof(P));
	P.d = d; P.N = N; P.M = M; P.a = a; P.k = k; P.v = v; P.shift = shift;
	P.volume = orc_volume(d, a); P.sigma = sigma; P.e0 = e0; P.rescale = 1.0; P.nthreads = nthreads;
	orc_problem_alloc(&P);
	memcpy(P.s, s, sizeof(double) * (size_t)(N * d));
	double E = orc_eval(&P, F_out != NULL);
	if (F_out) memcpy(F_out, P.F, sizeof(double) * (size_t)(N * d));
	orc_problem_free(&P);
	return E;
}

/* One RelaxStructure_NLOPT_inner pass (StructureOptimization.cpp:881-979, Switch 0) with the supplied
 * L-BFGS: skip if E < fmin; x = s*rescale (GetParam :710-776); minimize; s = wrap(x/rescale). */
static int orc_relax_pass(orc_problem *P, orc_plis_fn plis, double rescale, int maxstep, double fmin, int mf,
                          double *s_inout, int *evals_out)
{
	P->rescale = 1.0;
	memcpy(P->s, s_inout, sizeof(double) * (size_t)(P->N * P->d));
	double E0 = orc_eval(P, 0);
	if (E0 < fmin) return 0;
	int64_t n = P->N * P->d;
	double *x = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
	for (int64_t i = 0; i < n; i++) x[i] = s_inout[i] * rescale;
	P->rescale = rescale;
	double minf = 0;
	int evals = 0;
	int st = plis((int)n, orc_objective, P, x, &minf, fmin, maxstep, mf, &evals, 0);
	for (int64_t i = 0; i < n; i++) s_inout[i] = orc_wrap(x[i] / rescale);
	free(x);
	if (evals_out) *evals_out += evals;
	return st;
}

/* RelaxStructure_NLOPT_Emin (StructureOptimization.cpp:1073-1084) -- and RelaxStructure_NLOPT (:1046-1072)
 * when Emin = -MaxEnergy (-1e10): passes with rescale 10000, 1, 0.01, stopping between passes once
 * E < Emin.  Returns the final total energy; evals_out receives the objective evaluations of all passes. */
double orc_relax_emin(int d, int64_t N, double *s_inout, const double *a, int64_t M, const double *k, const double *v,
                      double shift, double sigma, double e0, int maxstep, double Emin, int mf, orc_plis_fn plis,
                      int nthreads, int *evals_out, int *status_out)
{
	orc_problem P;
	memset(&P, 0, sizeof(P));
	P.d = d; P.N = N; P.M = M; P.a = a; P.k = k; P.v = v; P.shift = shift;
	P.volume = orc_volume(d, a); P.sigma = sigma; P.e0 = e0; P.nthreads = nthreads;
	orc_problem_alloc(&P);
	const double rescales[3] = {10000, 1, 0.01};
	int evals = 0, st = 0;
	int emin_mode = Emin > -1e10;
	double E = 0;
	for (int p = 0; p < 3; p++) {
		st = orc_relax_pass(&P, plis, rescales[p], maxstep, Emin, mf, s_inout, &evals);
		P.rescale = 1.0;
		memcpy(P.s, s_inout, sizeof(double) * (size_t)(N * d));
		E = orc_eval(&P, 0);
		if (emin_mode && E < Emin) break;
	}
	if (evals_out) *evals_out = evals;
	if (status_out) *status_out = st;
	orc_problem_free(&P);
	return E;
}

int orc_max_threads(void) { return omp_get_max_threads(); }

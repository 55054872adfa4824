/* mmsp_oracle.c -- TEST INFRASTRUCTURE: CPU restatement of the reference's
 * explicit finite-difference update path, used only as the checker by the
 * tests, smoke() and the CPU-baseline leg of bench.py. Product code never
 * links or calls it.
 *
 * Every function follows the reference's evaluation order so that, with the
 * same cube routine, results are bitwise those of the reference build
 * (g++ -O3 without -march: no FMA contraction, libm pow for cubes).
 *
 *   check_boundary        include/MMSP.utility.cpp:42-57 (serial branch)
 *   node <-> coordinates  include/MMSP.grid.cpp:306-317, 375-411
 *   scalar laplacian      include/MMSP.grid.cpp:2049-2066
 *   vector laplacian      include/MMSP.grid.cpp:2068-2088
 *   Allen-Cahn update     examples/phase_transitions/allen-cahn/allen-cahn.cpp:58-73
 *   Cahn-Hilliard update  examples/phase_transitions/cahn-hilliard/explicit/cahn-hilliard.cpp:59-80
 *   grain growth update   examples/coarsening/grain_growth/isotropic/phase_field/graingrowth.cpp:84-110
 *   Zener pinning         examples/coarsening/zener_pinning/isotropic/phase_field/zener.cpp:104-141
 *   Ostwald ripening      examples/coarsening/ostwald_ripening/isotropic/phase_field/ostwald.cpp:56-123
 *   multi-phase GG        examples/coarsening/grain_growth/anisotropic/phase_field/graingrowth.cpp:94-191
 *   model A / model B     examples/phase_transitions/model_A/model_A.cpp:74-105,
 *                         examples/phase_transitions/model_B/model_B.cpp:74-111 (noise: normal
 *                         deviates as gaussian(), model_A.cpp:13-39, from Philox4x32-10 and
 *                         Box-Muller instead of rand() and the polar method)
 *
 * Cube: the reference evaluates pow(phi,3) with glibc pow (a call, checked in
 * the disassembly of the reference build). ORC_CUBE_POW reproduces that;
 * ORC_CUBE_CR is the correctly rounded cube (double-double product, one final
 * rounding) that the device kernels use. The two differ in ~1e-3 of inputs by
 * one ulp; the tests measure what that does to trajectories.
 *
 * Precision forms:
 *   *_f64      double storage and arithmetic (the reference with T = double)
 *   *_f32ref   float storage with the reference's mixed semantics for T = float
 *              (laplacian accumulates a double term into a float per axis;
 *              pow, dt and 2.0 promote the update expression to double)
 *   *_f32      float storage and float arithmetic with the same operation
 *              order (the fast FP32 device path; not a reference form)
 *
 * Reductions (mass, free energy) are summed with Neumaier compensation in
 * long double.
 *
 * Pinning: each function is checked by tests/test_oracle_*.py against golden
 * vectors written by the reference itself (oracle/_ref, tests/golden/) and
 * against closed forms (constant and Fourier-mode laplacians, the uniform-state
 * scalar map, fixed points, symmetry, mass invariance).
 */
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum { ORC_MIRROR = 0, ORC_NEUMANN = 1, ORC_PERIODIC = 2, ORC_PARALLEL = 3, ORC_DIRICHLET = 4 };
enum { ORC_CUBE_POW = 0, ORC_CUBE_CR = 1 };
enum { ORC_MU_POLY = 0, ORC_MU_DW = 1 };

typedef struct {
	int dim;
	int n[3];
	int b0[3];
	int b1[3];
	double dx[3];
} orc_grid;

/* ---- seeded input generator (workload definition) ---------------------- */
/* Element i of a seeded stream is the (i+1)-th output of SplitMix64 started
 * from state `seed` (Steele, Lea, Flood 2014; Vigna's reference constants). */
uint64_t orc_splitmix64(uint64_t seed, uint64_t i)
{
	uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

void orc_fill_uniform(double* out, int64_t n, uint64_t seed, double lo, double hi)
{
	for (int64_t i = 0; i < n; i++) {
		double u = (double)(orc_splitmix64(seed, (uint64_t)i) >> 11) * 0x1.0p-53;
		out[i] = lo + (hi - lo) * u;
	}
}

/* ---- boundary mapping and indexing -------------------------------------- */
int orc_check_boundary(int x, int x0, int x1, int b0, int b1)
{
	if (x < x0) {
		if (b0 == ORC_NEUMANN || b0 == ORC_DIRICHLET) x = x0;
		else if (b0 == ORC_PERIODIC) x = x1 - (x0 - x);
		else if (b0 == ORC_MIRROR) x = 2 * x0 - x;
	} else if (x >= x1) {
		if (b1 == ORC_NEUMANN || b1 == ORC_DIRICHLET) x = x1 - 1;
		else if (b1 == ORC_PERIODIC) x = x0 + (x - x1);
		else if (b1 == ORC_MIRROR) x = 2 * (x1 - 1) - x;
	}
	return x;
}

static int64_t nodes_of(const orc_grid* g)
{
	int64_t n = 1;
	for (int d = 0; d < g->dim; d++) n *= g->n[d];
	return n;
}

/* stride of axis d: product of the extents of the faster axes */
static int64_t stride_of(const orc_grid* g, int d)
{
	int64_t s = 1;
	for (int k = d + 1; k < g->dim; k++) s *= g->n[k];
	return s;
}

static void position(const orc_grid* g, int64_t i, int x[3])
{
	for (int d = 0; d < g->dim; d++) {
		int64_t s = stride_of(g, d);
		x[d] = (int)(i / s);
		i -= (int64_t)x[d] * s;
	}
}

/* node index of the neighbour of x shifted by `off` along axis d */
static int64_t neighbour(const orc_grid* g, const int x[3], int d, int off)
{
	int64_t idx = 0;
	for (int k = 0; k < g->dim; k++) {
		int xk = x[k];
		if (k == d) xk = orc_check_boundary(xk + off, 0, g->n[k], g->b0[k], g->b1[k]);
		idx += (int64_t)xk * stride_of(g, k);
	}
	return idx;
}

/* ---- cubes -------------------------------------------------------------- */
double orc_cube_cr(double x)
{
	double p = x * x;
	double e = fma(x, x, -p);     /* x*x = p + e exactly */
	double hi = p * x;
	double lo = fma(p, x, -hi);   /* p*x = hi + lo exactly */
	lo = lo + e * x;
	return hi + lo;
}

static double cube64(double x, int mode)
{
	return mode == ORC_CUBE_POW ? pow(x, 3.0) : orc_cube_cr(x);
}

/* ---- laplacians ---------------------------------------------------------- */
static double weight_of(const orc_grid* g, int d)
{
	return 1.0 / (g->dx[d] * g->dx[d]);
}

double orc_laplacian_f64(const orc_grid* g, const double* a, int64_t i)
{
	int x[3] = {0, 0, 0};
	position(g, i, x);
	double y = a[i];
	double lap = 0.0;
	for (int d = 0; d < g->dim; d++) {
		double yh = a[neighbour(g, x, d, +1)];
		double yl = a[neighbour(g, x, d, -1)];
		lap += weight_of(g, d) * (yh - 2.0 * y + yl);
	}
	return lap;
}

/* T = float in the reference: each axis adds a double term into a float. */
float orc_laplacian_f32ref(const orc_grid* g, const float* a, int64_t i)
{
	int x[3] = {0, 0, 0};
	position(g, i, x);
	float y = a[i];
	float lap = 0.0f;
	for (int d = 0; d < g->dim; d++) {
		float yh = a[neighbour(g, x, d, +1)];
		float yl = a[neighbour(g, x, d, -1)];
		lap = (float)((double)lap + weight_of(g, d) * ((double)yh - 2.0 * (double)y + (double)yl));
	}
	return lap;
}

float orc_laplacian_f32(const orc_grid* g, const float* a, int64_t i)
{
	int x[3] = {0, 0, 0};
	position(g, i, x);
	float y = a[i];
	float lap = 0.0f;
	for (int d = 0; d < g->dim; d++) {
		float yh = a[neighbour(g, x, d, +1)];
		float yl = a[neighbour(g, x, d, -1)];
		float w = (float)weight_of(g, d);
		lap += w * (yh - 2.0f * y + yl);
	}
	return lap;
}

/* Node loops of the scalar steppers run on OpenMP threads for grids of 2^16
 * nodes or more. Each node's arithmetic is unchanged, so the results do not
 * depend on the thread count (orc_set_threads). */
#define ORC_PAR_NODES _Pragma("omp parallel for schedule(static) if (n >= 65536)")

int orc_set_threads(int t)
{
#ifdef _OPENMP
	if (t > 0) omp_set_num_threads(t);
	return omp_get_max_threads();
#else
	(void)t;
	return 1;
#endif
}

/* ---- Allen-Cahn: p = {r, u, K, M, dt} ----------------------------------- */
/* new = phi - dt*M*(-r*phi + u*phi^3 - K*lap)   (allen-cahn.cpp:70) */
void orc_ac_steps_f64(const orc_grid* g, double* a, double* b, const double* p, int steps, int cube)
{
	const double r = p[0], u = p[1], K = p[2], M = p[3], dt = p[4];
	const int64_t n = nodes_of(g);
	for (int s = 0; s < steps; s++) {
		ORC_PAR_NODES
		for (int64_t i = 0; i < n; i++) {
			double phi = a[i];
			b[i] = phi - dt * M * (-r * phi + u * cube64(phi, cube) - K * orc_laplacian_f64(g, a, i));
		}
		memcpy(a, b, (size_t)n * sizeof(double));
	}
}

void orc_ac_steps_f32ref(const orc_grid* g, float* a, float* b, const double* p, int steps, int cube)
{
	const double r = p[0], u = p[1], K = p[2], M = p[3], dt = p[4];
	const int64_t n = nodes_of(g);
	for (int s = 0; s < steps; s++) {
		ORC_PAR_NODES
		for (int64_t i = 0; i < n; i++) {
			float phi = a[i];
			float lap = orc_laplacian_f32ref(g, a, i);
			b[i] = (float)((double)phi - dt * M * (-r * (double)phi + u * cube64((double)phi, cube) - K * (double)lap));
		}
		memcpy(a, b, (size_t)n * sizeof(float));
	}
}

void orc_ac_steps_f32(const orc_grid* g, float* a, float* b, const double* p, int steps)
{
	const float r = (float)p[0], u = (float)p[1], K = (float)p[2], M = (float)p[3], dt = (float)p[4];
	const int64_t n = nodes_of(g);
	for (int s = 0; s < steps; s++) {
		ORC_PAR_NODES
		for (int64_t i = 0; i < n; i++) {
			float phi = a[i];
			float c3 = phi * phi * phi;
			b[i] = phi - dt * M * (-r * phi + u * c3 - K * orc_laplacian_f32(g, a, i));
		}
		memcpy(a, b, (size_t)n * sizeof(float));
	}
}

/* ---- Cahn-Hilliard ------------------------------------------------------- */
/* ORC_MU_POLY: p = {r, u, K, M, dt}, mu = -r*phi + u*phi^3 - K*lap  (cahn-hilliard.cpp:71)
 * ORC_MU_DW:   p = {kappa, M, dt},  mu = 2c(1-c)(1-2c) - kappa*lap
 * then new = old + dt*M*lap(mu)                                      (cahn-hilliard.cpp:76) */
void orc_ch_steps_f64(const orc_grid* g, double* a, double* b, double* mu, const double* p, int form,
                      int steps, int cube)
{
	const int64_t n = nodes_of(g);
	for (int s = 0; s < steps; s++) {
		ORC_PAR_NODES
		for (int64_t i = 0; i < n; i++) {
			double c = a[i];
			double lap = orc_laplacian_f64(g, a, i);
			if (form == ORC_MU_POLY)
				mu[i] = -p[0] * c + p[1] * cube64(c, cube) - p[2] * lap;
			else
				mu[i] = 2.0 * c * (1.0 - c) * (1.0 - 2.0 * c) - p[0] * lap;
		}
		const double M = form == ORC_MU_POLY ? p[3] : p[1];
		const double dt = form == ORC_MU_POLY ? p[4] : p[2];
		ORC_PAR_NODES
		for (int64_t i = 0; i < n; i++)
			b[i] = a[i] + dt * M * orc_laplacian_f64(g, mu, i);
		memcpy(a, b, (size_t)n * sizeof(double));
	}
}

void orc_ch_steps_f32ref(const orc_grid* g, float* a, float* b, float* mu, const double* p, int form,
                         int steps, int cube)
{
	const int64_t n = nodes_of(g);
	for (int s = 0; s < steps; s++) {
		ORC_PAR_NODES
		for (int64_t i = 0; i < n; i++) {
			double c = a[i];
			double lap = orc_laplacian_f32ref(g, a, i);
			if (form == ORC_MU_POLY)
				mu[i] = (float)(-p[0] * c + p[1] * cube64(c, cube) - p[2] * lap);
			else
				mu[i] = (float)(2.0 * c * (1.0 - c) * (1.0 - 2.0 * c) - p[0] * lap);
		}
		const double M = form == ORC_MU_POLY ? p[3] : p[1];
		const double dt = form == ORC_MU_POLY ? p[4] : p[2];
		ORC_PAR_NODES
		for (int64_t i = 0; i < n; i++)
			b[i] = (float)((double)a[i] + dt * M * (double)orc_laplacian_f32ref(g, mu, i));
		memcpy(a, b, (size_t)n * sizeof(float));
	}
}

void orc_ch_steps_f32(const orc_grid* g, float* a, float* b, float* mu, const double* p, int form, int steps)
{
	const int64_t n = nodes_of(g);
	for (int s = 0; s < steps; s++) {
		ORC_PAR_NODES
		for (int64_t i = 0; i < n; i++) {
			float c = a[i];
			float lap = orc_laplacian_f32(g, a, i);
			if (form == ORC_MU_POLY)
				mu[i] = -(float)p[0] * c + (float)p[1] * (c * c * c) - (float)p[2] * lap;
			else
				mu[i] = 2.0f * c * (1.0f - c) * (1.0f - 2.0f * c) - (float)p[0] * lap;
		}
		const float M = (float)(form == ORC_MU_POLY ? p[3] : p[1]);
		const float dt = (float)(form == ORC_MU_POLY ? p[4] : p[2]);
		ORC_PAR_NODES
		for (int64_t i = 0; i < n; i++)
			b[i] = a[i] + dt * M * orc_laplacian_f32(g, mu, i);
		memcpy(a, b, (size_t)n * sizeof(float));
	}
}

/* ---- noise of the stochastic forms ---------------------------------------- */
/* Philox4x32-10 (Salmon, Moraes, Dror and Shaw, SC'11): ten rounds; in each,
 * words 0 and 2 are multiplied (32 x 32 -> 64 bits) by the round
 * multipliers, the high halves xored with words 1 and 3 and the key, and the
 * key advanced by the Weyl constants. Pinned by the known-answer vectors of
 * the Random123 distribution. */
void orc_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out)
{
	uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
	for (int r = 0; r < 10; r++) {
		const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
		const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
		c0 = n0;
		c1 = (uint32_t)p1;
		c2 = n2;
		c3 = (uint32_t)p0;
		k0 += 0x9E3779B9u;
		k1 += 0xBB67AE85u;
	}
	out[0] = c0;
	out[1] = c1;
	out[2] = c2;
	out[3] = c3;
}

/* ln u for u in (0, 1], a normal float, from +, -, *, / only: u = m 2^e, m in
 * (1/sqrt2, sqrt2], ln m = 2 atanh s, s = (m-1)/(m+1), series to s^9. */
float orc_ln_unit(float u)
{
	uint32_t bits;
	memcpy(&bits, &u, 4);
	int e = (int)(bits >> 23) - 127;
	const uint32_t mb = (bits & 0x7FFFFFu) | 0x3F800000u;
	float m;
	memcpy(&m, &mb, 4);
	if (m > 1.41421356f) {
		m = m * 0.5f;
		e += 1;
	}
	const float s = (m - 1.0f) / (m + 1.0f), z = s * s;
	float q = 1.0f / 9.0f;
	q = 1.0f / 7.0f + z * q;
	q = 1.0f / 5.0f + z * q;
	q = 1.0f / 3.0f + z * q;
	q = 1.0f + z * q;
	return (float)e * 0.693147182f + 2.0f * s * q;
}

/* cos(2 pi t) and sin(2 pi t), t in [0, 1) a multiple of 2^-24: quadrant q and
 * f = 4t - q exact; theta = f pi/2 folded into [0, pi/4]; Taylor series to
 * theta^10 (cos) and theta^9 (sin) give A = cos(f pi/2), B = sin(f pi/2). */
void orc_cos_sin_2pi(float t, float* cosv, float* sinv)
{
	const float v = t * 4.0f;
	const int q = (int)v;
	const float f = v - (float)q;
	const int hi = f > 0.5f;
	const float h = hi ? 1.0f - f : f;
	const float th = h * 1.57079633f, z = th * th;
	const float cs = 1.0f + z * (-1.0f / 2.0f + z * (1.0f / 24.0f + z * (-1.0f / 720.0f +
	                 z * (1.0f / 40320.0f + z * (-1.0f / 3628800.0f)))));
	const float sn = th * (1.0f + z * (-1.0f / 6.0f + z * (1.0f / 120.0f + z * (-1.0f / 5040.0f +
	                 z * (1.0f / 362880.0f)))));
	const float A = hi ? sn : cs, B = hi ? cs : sn;
	*cosv = q == 0 ? A : q == 1 ? -B : q == 2 ? -A : B;
	*sinv = q == 0 ? B : q == 1 ? A : q == 2 ? -B : -A;
}

/* Standard normal deviate of the node with axis-0 index x and index r within
 * its axis-0 plane: Box-Muller on the first two words of Philox4x32-10 with key
 * {seed lo, seed hi} and counter {r, x >> 1, step lo, step hi}: u1 = (w0 + 1/2)
 * 2^-32, u2 = (w1 >> 8) 2^-24, R = sqrt(-2 ln u1); R cos(2 pi u2) for even x,
 * R sin(2 pi u2) for odd x, in float. (The reference's gaussian(),
 * model_A.cpp:13-39, uses the polar method on rand() and likewise hands out the
 * two values of each pair in turn: the same distribution.) */
float orc_normal(uint64_t seed, uint64_t step, uint64_t x, uint64_t r)
{
	const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
	const uint32_t ctr[4] = {(uint32_t)r, (uint32_t)(x >> 1), (uint32_t)step, (uint32_t)(step >> 32)};
	uint32_t w[4];
	orc_philox4x32_10(ctr, key, w);
	const float u1 = ((float)w[0] + 0.5f) * 0x1p-32f;
	const float u2 = (float)(w[1] >> 8) * 0x1p-24f;
	float c, sn;
	orc_cos_sin_2pi(u2, &c, &sn);
	const float R = sqrtf(-2.0f * orc_ln_unit(u1));
	return (x & 1) ? R * sn : R * c;
}

/* the deviate of node i: x = i / plane, r = i mod plane (plane = nodes per axis-0 index) */
static float node_normal(const orc_grid* g, uint64_t seed, uint64_t step, int64_t i)
{
	const int64_t plane = nodes_of(g) / g->n[0];
	return orc_normal(seed, step, (uint64_t)(i / plane), (uint64_t)(i % plane));
}

/* ---- model A: p = {r, u, K, M, dt, sigma, seed}; step0 = the state's step count
 * new = phi - dt*M*(-r*phi + u*phi^3 - K*lap + noise), noise = 0 + g*sigma
 * (model_A.cpp:98-100) */
void orc_ac_noise_steps_f64(const orc_grid* g, double* a, double* b, const double* p, int steps, uint64_t step0,
                            int cube)
{
	const double r = p[0], u = p[1], K = p[2], M = p[3], dt = p[4], sigma = p[5];
	const uint64_t seed = (uint64_t)p[6];
	const int64_t n = nodes_of(g);
	for (int s = 0; s < steps; s++) {
		ORC_PAR_NODES
		for (int64_t i = 0; i < n; i++) {
			double phi = a[i];
			double noise = 0.0 + (double)node_normal(g, seed, step0 + (uint64_t)s, i) * sigma;
			b[i] = phi - dt * M * (-r * phi + u * cube64(phi, cube) - K * orc_laplacian_f64(g, a, i) + noise);
		}
		memcpy(a, b, (size_t)n * sizeof(double));
	}
}

/* T = float: the deviate is rounded to float (T noise) and promoted again */
void orc_ac_noise_steps_f32ref(const orc_grid* g, float* a, float* b, const double* p, int steps, uint64_t step0,
                               int cube)
{
	const double r = p[0], u = p[1], K = p[2], M = p[3], dt = p[4], sigma = p[5];
	const uint64_t seed = (uint64_t)p[6];
	const int64_t n = nodes_of(g);
	for (int s = 0; s < steps; s++) {
		ORC_PAR_NODES
		for (int64_t i = 0; i < n; i++) {
			float phi = a[i];
			float lap = orc_laplacian_f32ref(g, a, i);
			float noise = (float)(0.0 + (double)node_normal(g, seed, step0 + (uint64_t)s, i) * sigma);
			b[i] = (float)((double)phi -
			               dt * M * (-r * (double)phi + u * cube64((double)phi, cube) - K * (double)lap + (double)noise));
		}
		memcpy(a, b, (size_t)n * sizeof(float));
	}
}

void orc_ac_noise_steps_f32(const orc_grid* g, float* a, float* b, const double* p, int steps, uint64_t step0)
{
	const float r = (float)p[0], u = (float)p[1], K = (float)p[2], M = (float)p[3], dt = (float)p[4];
	const float sigma = (float)p[5];
	const uint64_t seed = (uint64_t)p[6];
	const int64_t n = nodes_of(g);
	for (int s = 0; s < steps; s++) {
		ORC_PAR_NODES
		for (int64_t i = 0; i < n; i++) {
			float phi = a[i];
			float c3 = phi * phi * phi;
			float noise = node_normal(g, seed, step0 + (uint64_t)s, i) * sigma;
			b[i] = phi - dt * M * (-r * phi + u * c3 - K * orc_laplacian_f32(g, a, i) + noise);
		}
		memcpy(a, b, (size_t)n * sizeof(float));
	}
}

/* ---- model B: p = {r, u, K, M, dt, sigma, seed}
 * mu = -r*phi + u*phi^3 - K*lap + noise (model_B.cpp:99-101); new = old + dt*M*lap(mu) (:105-106).
 * spinodal.cpp:94-102 is r = u = K = 1, M = 1/(2 dV), sigma = sqrt(eps dt/dV). */
void orc_ch_noise_steps_f64(const orc_grid* g, double* a, double* b, double* mu, const double* p, int steps,
                            uint64_t step0, int cube)
{
	const double sigma = p[5];
	const uint64_t seed = (uint64_t)p[6];
	const int64_t n = nodes_of(g);
	for (int s = 0; s < steps; s++) {
		ORC_PAR_NODES
		for (int64_t i = 0; i < n; i++) {
			double c = a[i];
			double noise = 0.0 + (double)node_normal(g, seed, step0 + (uint64_t)s, i) * sigma;
			mu[i] = -p[0] * c + p[1] * cube64(c, cube) - p[2] * orc_laplacian_f64(g, a, i) + noise;
		}
		ORC_PAR_NODES
		for (int64_t i = 0; i < n; i++) b[i] = a[i] + p[4] * p[3] * orc_laplacian_f64(g, mu, i);
		memcpy(a, b, (size_t)n * sizeof(double));
	}
}

void orc_ch_noise_steps_f32ref(const orc_grid* g, float* a, float* b, float* mu, const double* p, int steps,
                               uint64_t step0, int cube)
{
	const double sigma = p[5];
	const uint64_t seed = (uint64_t)p[6];
	const int64_t n = nodes_of(g);
	for (int s = 0; s < steps; s++) {
		ORC_PAR_NODES
		for (int64_t i = 0; i < n; i++) {
			double c = a[i];
			double lap = orc_laplacian_f32ref(g, a, i);
			float noise = (float)(0.0 + (double)node_normal(g, seed, step0 + (uint64_t)s, i) * sigma);
			mu[i] = (float)(-p[0] * c + p[1] * cube64(c, cube) - p[2] * lap + (double)noise);
		}
		ORC_PAR_NODES
		for (int64_t i = 0; i < n; i++)
			b[i] = (float)((double)a[i] + p[4] * p[3] * (double)orc_laplacian_f32ref(g, mu, i));
		memcpy(a, b, (size_t)n * sizeof(float));
	}
}

void orc_ch_noise_steps_f32(const orc_grid* g, float* a, float* b, float* mu, const double* p, int steps,
                            uint64_t step0)
{
	const float sigma = (float)p[5];
	const uint64_t seed = (uint64_t)p[6];
	const int64_t n = nodes_of(g);
	for (int s = 0; s < steps; s++) {
		ORC_PAR_NODES
		for (int64_t i = 0; i < n; i++) {
			float c = a[i];
			float lap = orc_laplacian_f32(g, a, i);
			float noise = node_normal(g, seed, step0 + (uint64_t)s, i) * sigma;
			mu[i] = -(float)p[0] * c + (float)p[1] * (c * c * c) - (float)p[2] * lap + noise;
		}
		const float M = (float)p[3], dt = (float)p[4];
		ORC_PAR_NODES
		for (int64_t i = 0; i < n; i++) b[i] = a[i] + dt * M * orc_laplacian_f32(g, mu, i);
		memcpy(a, b, (size_t)n * sizeof(float));
	}
}

/* ---- grain growth: fields node-major (Q values per node), p = {dt} ------- */
/* lap_j as MMSP.grid.cpp:2068-2088; sum = sum_j phi_j^2 in T;
 * new_j = phi_j - dt*(-phi_j - phi_j^3 + 2.0*(phi_j*sum - lap_j))   (graingrowth.cpp:98-105) */
static void gg_lap_f64(const orc_grid* g, int Q, const double* a, int64_t i, double* lap)
{
	int x[3] = {0, 0, 0};
	position(g, i, x);
	const double* y = a + i * Q;
	for (int j = 0; j < Q; j++) lap[j] = 0.0;
	for (int d = 0; d < g->dim; d++) {
		const double* yh = a + neighbour(g, x, d, +1) * Q;
		const double* yl = a + neighbour(g, x, d, -1) * Q;
		double w = weight_of(g, d);
		for (int j = 0; j < Q; j++) lap[j] += w * (yh[j] - 2.0 * y[j] + yl[j]);
	}
}

void orc_gg_steps_f64(const orc_grid* g, int Q, double* a, double* b, double dt, int steps, int cube)
{
	const int64_t n = nodes_of(g);
	for (int s = 0; s < steps; s++) {
#pragma omp parallel if (n >= 65536)
		{
			double* lap = (double*)malloc(sizeof(double) * (size_t)Q);
#pragma omp for schedule(static)
			for (int64_t i = 0; i < n; i++) {
				const double* phi = a + i * Q;
				gg_lap_f64(g, Q, a, i, lap);
				double sum = 0.0;
				for (int j = 0; j < Q; j++) sum += phi[j] * phi[j];
				for (int j = 0; j < Q; j++)
					b[i * Q + j] = phi[j] - dt * (-phi[j] - cube64(phi[j], cube) + 2.0 * (phi[j] * sum - lap[j]));
			}
			free(lap);
		}
		memcpy(a, b, (size_t)n * (size_t)Q * sizeof(double));
	}
}

void orc_gg_steps_f32ref(const orc_grid* g, int Q, float* a, float* b, double dt, int steps, int cube)
{
	const int64_t n = nodes_of(g);
	float* lap = (float*)malloc(sizeof(float) * (size_t)Q);
	int x[3] = {0, 0, 0};
	for (int s = 0; s < steps; s++) {
		for (int64_t i = 0; i < n; i++) {
			const float* phi = a + i * Q;
			position(g, i, x);
			for (int j = 0; j < Q; j++) lap[j] = 0.0f;
			for (int d = 0; d < g->dim; d++) {
				const float* yh = a + neighbour(g, x, d, +1) * Q;
				const float* yl = a + neighbour(g, x, d, -1) * Q;
				double w = weight_of(g, d);
				for (int j = 0; j < Q; j++)
					lap[j] = (float)((double)lap[j] + w * ((double)yh[j] - 2.0 * (double)phi[j] + (double)yl[j]));
			}
			float sum = 0.0f;
			for (int j = 0; j < Q; j++) sum += phi[j] * phi[j];
			for (int j = 0; j < Q; j++) {
				double pj = phi[j];
				float ps = phi[j] * sum;
				float inner = ps - lap[j];
				b[i * Q + j] = (float)(pj - dt * (-pj - cube64(pj, cube) + 2.0 * (double)inner));
			}
		}
		memcpy(a, b, (size_t)n * (size_t)Q * sizeof(float));
	}
	free(lap);
}

void orc_gg_steps_f32(const orc_grid* g, int Q, float* a, float* b, double dt64, int steps)
{
	const int64_t n = nodes_of(g);
	const float dt = (float)dt64;
	for (int s = 0; s < steps; s++) {
#pragma omp parallel if (n >= 65536)
		{
			float* lap = (float*)malloc(sizeof(float) * (size_t)Q);
			int x[3] = {0, 0, 0};
#pragma omp for schedule(static)
			for (int64_t i = 0; i < n; i++) {
				const float* phi = a + i * Q;
				position(g, i, x);
				for (int j = 0; j < Q; j++) lap[j] = 0.0f;
				for (int d = 0; d < g->dim; d++) {
					const float* yh = a + neighbour(g, x, d, +1) * Q;
					const float* yl = a + neighbour(g, x, d, -1) * Q;
					float w = (float)weight_of(g, d);
					for (int j = 0; j < Q; j++) lap[j] += w * (yh[j] - 2.0f * phi[j] + yl[j]);
				}
				float sum = 0.0f;
				for (int j = 0; j < Q; j++) sum += phi[j] * phi[j];
				for (int j = 0; j < Q; j++) {
					float pj = phi[j];
					b[i * Q + j] = pj - dt * (-pj - pj * pj * pj + 2.0f * (pj * sum - lap[j]));
				}
			}
			free(lap);
		}
		memcpy(a, b, (size_t)n * (size_t)Q * sizeof(float));
	}
}

/* ---- Zener pinning: zener.cpp:117-135, p = {dt} -------------------------- */
/* grain growth with `double sum` and field 0 (the particles) held */
void orc_zener_steps_f64(const orc_grid* g, int Q, double* a, double* b, double dt, int steps, int cube)
{
	const int64_t n = nodes_of(g);
	double* lap = (double*)malloc(sizeof(double) * (size_t)Q);
	for (int s = 0; s < steps; s++) {
		for (int64_t i = 0; i < n; i++) {
			const double* phi = a + i * Q;
			gg_lap_f64(g, Q, a, i, lap);
			double sum = 0.0;
			for (int j = 0; j < Q; j++) sum += phi[j] * phi[j];
			for (int j = 0; j < Q; j++)
				b[i * Q + j] = j == 0 ? phi[j]
				                      : phi[j] - dt * (-phi[j] - cube64(phi[j], cube) + 2.0 * (phi[j] * sum - lap[j]));
		}
		memcpy(a, b, (size_t)n * (size_t)Q * sizeof(double));
	}
	free(lap);
}

/* the float form: float arithmetic in the same order */
void orc_zener_steps_f32(const orc_grid* g, int Q, float* a, float* b, double dt64, int steps)
{
	const int64_t n = nodes_of(g);
	const float dt = (float)dt64;
	float* lap = (float*)malloc(sizeof(float) * (size_t)Q);
	int x[3] = {0, 0, 0};
	for (int s = 0; s < steps; s++) {
		for (int64_t i = 0; i < n; i++) {
			const float* phi = a + i * Q;
			position(g, i, x);
			for (int j = 0; j < Q; j++) lap[j] = 0.0f;
			for (int d = 0; d < g->dim; d++) {
				const float* yh = a + neighbour(g, x, d, +1) * Q;
				const float* yl = a + neighbour(g, x, d, -1) * Q;
				float w = (float)weight_of(g, d);
				for (int j = 0; j < Q; j++) lap[j] += w * (yh[j] - 2.0f * phi[j] + yl[j]);
			}
			float sum = 0.0f;
			for (int j = 0; j < Q; j++) sum += phi[j] * phi[j];
			for (int j = 0; j < Q; j++) {
				float pj = phi[j];
				b[i * Q + j] = j == 0 ? pj : pj - dt * (-pj - pj * pj * pj + 2.0f * (pj * sum - lap[j]));
			}
		}
		memcpy(a, b, (size_t)n * (size_t)Q * sizeof(float));
	}
	free(lap);
}

/* ---- Ostwald ripening: ostwald.cpp:68-122 ----------------------------------
 * p = {dt, L, D, Calpha, Cbeta, A, delta, epsilon, kappa}; field 0 = C.
 * ORC_CUBE_POW evaluates pow(x,3) and pow(x,2) with libm as the reference does;
 * ORC_CUBE_CR uses the correctly rounded cube and x*x, as the device does. */
static double sq64(double x, int mode) { return mode == ORC_CUBE_POW ? pow(x, 2.0) : x * x; }

void orc_ostwald_steps_f64(const orc_grid* g, int Q, double* a, double* b, double* w, const double* p, int steps,
                           int cube)
{
	const double dt = p[0], L = p[1], D = p[2], Calpha = p[3], Cbeta = p[4], A = p[5], delta = p[6];
	const double epsilon = p[7], kappa = p[8];
	const double Cmatrix = 0.5 * (Calpha + Cbeta);
	const double B = A / pow(Cbeta - Cmatrix, 2);
	const double gamma = 2.0 / pow(Cbeta - Calpha, 2);
	const double Dalpha = gamma / pow(delta, 2);
	const double Dbeta = gamma / pow(delta, 2);
	const int64_t n = nodes_of(g);
	double* vlap = (double*)malloc(sizeof(double) * (size_t)Q);
	orc_grid g1 = *g;
	for (int s = 0; s < steps; s++) {
		for (int64_t i = 0; i < n; i++) {
			double sum = 0.0;
			for (int j = 1; j < Q; j++) sum += sq64(a[i * Q + j], cube);
			const double C = a[i * Q];
			gg_lap_f64(g, Q, a, i, vlap); /* field 0's laplacian: the single-field form, same terms */
			const double lap = vlap[0];
			w[i] = -A * (C - Cmatrix) + B * cube64(C - Cmatrix, cube) + Dalpha * cube64(C - Calpha, cube) +
			       Dbeta * cube64(C - Cbeta, cube) - gamma * (C - Calpha) * sum - kappa * lap;
		}
		for (int64_t i = 0; i < n; i++) {
			const double lap = orc_laplacian_f64(&g1, w, i);
			const double C = a[i * Q];
			b[i * Q] = C + dt * D * lap;
			double sum = 0.0;
			for (int j = 1; j < Q; j++) sum += sq64(a[i * Q + j], cube);
			gg_lap_f64(g, Q, a, i, vlap);
			for (int j = 1; j < Q; j++) {
				const double value = a[i * Q + j];
				b[i * Q + j] = value - dt * L * (-gamma * sq64(C - Calpha, cube) + delta * cube64(value, cube) +
				                                 epsilon * value * (sum - sq64(value, cube)) - kappa * vlap[j]);
			}
		}
		memcpy(a, b, (size_t)n * (size_t)Q * sizeof(double));
	}
	free(vlap);
}

/* the float form: constants derived in double and rounded, float arithmetic */
void orc_ostwald_steps_f32(const orc_grid* g, int Q, float* a, float* b, float* w, const double* p, int steps)
{
	const double Cm64 = 0.5 * (p[3] + p[4]);
	const double gamma64 = 2.0 / pow(p[4] - p[3], 2);
	const float A = (float)p[5], B = (float)(p[5] / pow(p[4] - Cm64, 2)), gamma = (float)gamma64;
	const float Dalpha = (float)(gamma64 / pow(p[6], 2)), Dbeta = Dalpha, delta = (float)p[6];
	const float epsilon = (float)p[7], kappa = (float)p[8], Calpha = (float)p[3], Cbeta = (float)p[4];
	const float Cmatrix = (float)Cm64, dtD = (float)(p[0] * p[2]), dtL = (float)(p[0] * p[1]);
	const int64_t n = nodes_of(g);
	float* lap = (float*)malloc(sizeof(float) * (size_t)Q);
	int x[3] = {0, 0, 0};
	for (int s = 0; s < steps; s++) {
		for (int64_t i = 0; i < n; i++) {
			const float* phi = a + i * Q;
			position(g, i, x);
			for (int j = 0; j < Q; j++) lap[j] = 0.0f;
			for (int d = 0; d < g->dim; d++) {
				const float* yh = a + neighbour(g, x, d, +1) * Q;
				const float* yl = a + neighbour(g, x, d, -1) * Q;
				float wt = (float)weight_of(g, d);
				for (int j = 0; j < Q; j++) lap[j] += wt * (yh[j] - 2.0f * phi[j] + yl[j]);
			}
			float sum = 0.0f;
			for (int j = 1; j < Q; j++) sum += phi[j] * phi[j];
			const float C = phi[0];
			const float ca = C - Cmatrix, cb = C - Calpha, cc = C - Cbeta;
			w[i] = -A * ca + B * (ca * ca * ca) + Dalpha * (cb * cb * cb) + Dbeta * (cc * cc * cc) - gamma * cb * sum -
			       kappa * lap[0];
			for (int j = 1; j < Q; j++) {
				const float v = phi[j];
				b[i * Q + j] = v - dtL * (-gamma * (cb * cb) + delta * (v * v * v) + epsilon * v * (sum - v * v) -
				                          kappa * lap[j]);
			}
		}
		for (int64_t i = 0; i < n; i++) b[i * Q] = a[i * Q] + dtD * orc_laplacian_f32(g, w, i);
		memcpy(a, b, (size_t)n * (size_t)Q * sizeof(float));
	}
	free(lap);
}

/* ---- anisotropic (multi-phase) grain growth ---------------------------------
 * grain_growth/anisotropic/phase_field/graingrowth.cpp:94-191, p = {dt, width,
 * gamma, mu}; energy() and mobility() (anisotropy.hpp) return float. No pow:
 * the restatement is bitwise the reference. */
#define ORC_GGA_MAXQ 64
void orc_gga_steps_f64(const orc_grid* g, int Q, double* a, double* b, const double* p, int steps)
{
	const double dt = p[0], width = p[1];
	const double gamma = (double)(float)p[2], mu = (double)(float)p[3];
	const double eps = 4.0 / acos(-1.0) * sqrt(0.5 * gamma * width);
	const double w = 4.0 * gamma / width;
	const int64_t n = nodes_of(g);
	double lap[ORC_GGA_MAXQ], dFdp[ORC_GGA_MAXQ], dpdt[ORC_GGA_MAXQ];
	int s[ORC_GGA_MAXQ];
	int x[3] = {0, 0, 0};
	for (int st = 0; st < steps; st++) {
		for (int64_t i = 0; i < n; i++) {
			const double* phi = a + i * Q;
			position(g, i, x);
			double S = 0.0;
			for (int h = 0; h < Q; h++) {
				s[h] = 0;
				for (int j = 0; j < g->dim && !s[h]; j++)
					for (int k = -1; k <= 1 && !s[h]; k++)
						if (a[neighbour(g, x, j, k) * Q + h] > 0.0) s[h] = 1;
				S += s[h];
			}
			if (S < 2.0) {
				for (int h = 0; h < Q; h++) b[i * Q + h] = phi[h];
				continue;
			}
			gg_lap_f64(g, Q, a, i, lap);
			for (int h = 0; h < Q; h++) dFdp[h] = dpdt[h] = 0.0;
			for (int h = 0; h < Q; h++)
				if (s[h] > 0.0)
					for (int j = h + 1; j < Q; j++)
						if (s[j] > 0.0) {
							dFdp[h] += 0.5 * eps * eps * lap[j] + w * phi[j];
							dFdp[j] += 0.5 * eps * eps * lap[h] + w * phi[h];
							for (int k = j + 1; k < Q; k++)
								if (s[k] > 0.0) {
									dFdp[h] += 3.0 * phi[j] * phi[k];
									dFdp[j] += 3.0 * phi[k] * phi[h];
									dFdp[k] += 3.0 * phi[h] * phi[j];
								}
						}
			for (int h = 0; h < Q; h++)
				if (s[h] > 0.0)
					for (int j = h + 1; j < Q; j++)
						if (s[j] > 0.0) {
							dpdt[h] -= mu * (dFdp[h] - dFdp[j]);
							dpdt[j] -= mu * (dFdp[j] - dFdp[h]);
						}
			double sum = 0.0;
			for (int h = 0; h < Q; h++) {
				double value = phi[h] + dt * (2.0 / S) * dpdt[h];
				if (value > 1.0) value = 1.0;
				if (value < 0.0) value = 0.0;
				b[i * Q + h] = value;
				sum += value;
			}
			double rsum = 0.0;
			if (fabs(sum) > 0.0) rsum = 1.0 / sum;
			for (int h = 0; h < Q; h++) b[i * Q + h] *= rsum;
		}
		memcpy(a, b, (size_t)n * (size_t)Q * sizeof(double));
	}
}

/* the float form: constants formed in double and rounded, float arithmetic */
void orc_gga_steps_f32(const orc_grid* g, int Q, float* a, float* b, const double* p, int steps)
{
	const double gamma64 = (double)(float)p[2];
	const double eps64 = 4.0 / acos(-1.0) * sqrt(0.5 * gamma64 * p[1]);
	const float c2 = (float)(0.5 * eps64 * eps64), w = (float)(4.0 * gamma64 / p[1]);
	const float mu = (float)(double)(float)p[3], dt = (float)p[0];
	const int64_t n = nodes_of(g);
	float lap[ORC_GGA_MAXQ], dF[ORC_GGA_MAXQ], dp[ORC_GGA_MAXQ];
	int s[ORC_GGA_MAXQ];
	int x[3] = {0, 0, 0};
	for (int st = 0; st < steps; st++) {
		for (int64_t i = 0; i < n; i++) {
			const float* phi = a + i * Q;
			position(g, i, x);
			float S = 0.0f;
			for (int h = 0; h < Q; h++) {
				s[h] = 0;
				for (int j = 0; j < g->dim && !s[h]; j++)
					for (int k = -1; k <= 1 && !s[h]; k++)
						if (a[neighbour(g, x, j, k) * Q + h] > 0.0f) s[h] = 1;
				S += (float)s[h];
			}
			if (S < 2.0f) {
				for (int h = 0; h < Q; h++) b[i * Q + h] = phi[h];
				continue;
			}
			for (int h = 0; h < Q; h++) {
				lap[h] = 0.0f;
				dF[h] = dp[h] = 0.0f;
			}
			for (int d = 0; d < g->dim; d++) {
				const float* yh = a + neighbour(g, x, d, +1) * Q;
				const float* yl = a + neighbour(g, x, d, -1) * Q;
				const float wt = (float)weight_of(g, d);
				for (int h = 0; h < Q; h++)
					if (s[h]) lap[h] += wt * (yh[h] - 2.0f * phi[h] + yl[h]);
			}
			for (int h = 0; h < Q; h++)
				if (s[h])
					for (int j = h + 1; j < Q; j++)
						if (s[j]) {
							dF[h] += c2 * lap[j] + w * phi[j];
							dF[j] += c2 * lap[h] + w * phi[h];
							for (int k = j + 1; k < Q; k++)
								if (s[k]) {
									dF[h] += 3.0f * phi[j] * phi[k];
									dF[j] += 3.0f * phi[k] * phi[h];
									dF[k] += 3.0f * phi[h] * phi[j];
								}
						}
			for (int h = 0; h < Q; h++)
				if (s[h])
					for (int j = h + 1; j < Q; j++)
						if (s[j]) {
							dp[h] -= mu * (dF[h] - dF[j]);
							dp[j] -= mu * (dF[j] - dF[h]);
						}
			float sum = 0.0f;
			const float rate = dt * (2.0f / S);
			for (int h = 0; h < Q; h++) {
				float v = phi[h] + rate * dp[h];
				if (v > 1.0f) v = 1.0f;
				if (v < 0.0f) v = 0.0f;
				b[i * Q + h] = v;
				sum += v;
			}
			const float rsum = fabsf(sum) > 0.0f ? 1.0f / sum : 0.0f;
			for (int h = 0; h < Q; h++) b[i * Q + h] *= rsum;
		}
		memcpy(a, b, (size_t)n * (size_t)Q * sizeof(float));
	}
}

/* ---- reductions ----------------------------------------------------------- */
typedef struct {
	long double s, c;
} nsum;

static void nadd(nsum* a, long double v)
{
	long double t = a->s + v;
	if (fabsl(a->s) >= fabsl(v)) a->c += (a->s - t) + v;
	else a->c += (v - t) + a->s;
	a->s = t;
}

/* per-field mass sum_i a[i*Q+j] (node-major); float inputs via the _f32 entry */
void orc_mass_f64(const double* a, int64_t n, int Q, double* out)
{
	for (int j = 0; j < Q; j++) {
		nsum acc = {0.0L, 0.0L};
		for (int64_t i = 0; i < n; i++) nadd(&acc, a[i * Q + j]);
		out[j] = (double)(acc.s + acc.c);
	}
}

void orc_mass_f32(const float* a, int64_t n, int Q, double* out)
{
	for (int j = 0; j < Q; j++) {
		nsum acc = {0.0L, 0.0L};
		for (int64_t i = 0; i < n; i++) nadd(&acc, a[i * Q + j]);
		out[j] = (double)(acc.s + acc.c);
	}
}

/* Free energies, node terms evaluated in double, summed compensated.
 * Forward differences (the discrete gradient whose adjoint gives the
 * 5/7-point laplacian), neighbours through check_boundary.
 *   kind 0 (AC / CH poly, p = {r, u, K}):  -r/2 c^2 + u/4 c^4 + K/2 |grad+ c|^2
 *   kind 1 (CH double well, p = {kappa}):   c^2 (1-c)^2 + kappa/2 |grad+ c|^2
 *   kind 2 (grain growth, p = {kappa}):     sum_j (-phi_j^2/2 + phi_j^4/4)
 *                                           + sum_{j<k} phi_j^2 phi_k^2 + kappa/2 sum_j |grad+ phi_j|^2 */
static double bulk_term(int kind, const double* p, const double* v, int Q)
{
	if (kind == 0) {
		double c = v[0], c2 = c * c;
		return -0.5 * p[0] * c2 + 0.25 * p[1] * (c2 * c2);
	}
	if (kind == 1) {
		double c = v[0], m = 1.0 - c;
		return (c * c) * (m * m);
	}
	double e = 0.0, s2 = 0.0;
	for (int j = 0; j < Q; j++) {
		double q = v[j] * v[j];
		e += -0.5 * q + 0.25 * (q * q);
		e += q * s2; /* pairs j<k counted once: phi_j^2 * sum_{k<j} phi_k^2 */
		s2 += q;
	}
	return e;
}

static double energy_impl(const orc_grid* g, int Q, const double* a, int kind, const double* p)
{
	const int64_t n = nodes_of(g);
	const double kap = kind == 0 ? p[2] : p[0];
	nsum acc = {0.0L, 0.0L};
	int x[3] = {0, 0, 0};
	for (int64_t i = 0; i < n; i++) {
		position(g, i, x);
		double e = bulk_term(kind, p, a + i * Q, Q);
		double gsum = 0.0;
		for (int d = 0; d < g->dim; d++) {
			const double* yh = a + neighbour(g, x, d, +1) * Q;
			const double* y = a + i * Q;
			for (int j = 0; j < Q; j++) {
				double df = (yh[j] - y[j]) / g->dx[d];
				gsum += df * df;
			}
		}
		e += 0.5 * kap * gsum;
		nadd(&acc, e);
	}
	double vol = 1.0;
	for (int d = 0; d < g->dim; d++) vol *= g->dx[d];
	return (double)((acc.s + acc.c) * vol);
}

double orc_energy_f64(const orc_grid* g, int Q, const double* a, int kind, const double* p)
{
	return energy_impl(g, Q, a, kind, p);
}

double orc_energy_f32(const orc_grid* g, int Q, const float* a, int kind, const double* p)
{
	int64_t n = nodes_of(g) * Q;
	double* tmp = (double*)malloc(sizeof(double) * (size_t)n);
	for (int64_t i = 0; i < n; i++) tmp[i] = a[i];
	double e = energy_impl(g, Q, tmp, kind, p);
	free(tmp);
	return e;
}

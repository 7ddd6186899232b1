/*
 * Synthetic single-particle diffraction data for benchmarking and testing.
 *
 * Object: n_blobs 3-D Gaussian blobs, centres uniform in a ball of radius
 * blob_radius around the grid centre, widths uniform in [sigma_min,
 * sigma_max], peak amplitudes uniform in [0.5, 1].  Its Fourier transform is
 * analytic and separable per blob,
 *     F(k) = sum_j a_j (2 pi s_j^2)^(3/2) exp(-2 pi^2 s_j^2 |k|^2 / S^2)
 *                  exp(-2 pi i k.c_j / S),
 * so the intensity |F|^2 is evaluated exactly on the model grid instead of by
 * an FFT.  Voxel v holds frequency k = v - S/2 - 0.5 per axis: the reference
 * maps a pixel at coordinate c to voxel round(S/2 + 0.5 + c) (emc.c:385-387,
 * emc_cuda.cu:86-88), so that is where its model is centred.
 *
 * All random numbers come from a counter-based generator (a hash of seed,
 * stream and two counters), so results do not depend on the number of
 * threads or on the order of evaluation.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "emc_host.h"

enum { ST_BLOB = 1, ST_ORIENT = 2, ST_PHI = 3, ST_COUNT = 4, ST_MODEL = 5 };

static inline uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

double emc_uniform(uint64_t seed, uint64_t stream, uint64_t a, uint64_t b) {
  uint64_t h = mix64(seed);
  h = mix64(h ^ (stream * 0xd1b54a32d192ed03ULL));
  h = mix64(h ^ a);
  h = mix64(h ^ (b * 0x8cb92ba72f3d8dd7ULL));
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

void emc_quaternion_from_uniforms(double u1, double u2, double u3, float q[4]) {
  /* emc.c:27-38 */
  q[0] = (float)(sqrt(1 - u1) * sin(2.0 * M_PI * u2));
  q[1] = (float)(sqrt(1 - u1) * cos(2.0 * M_PI * u2));
  q[2] = (float)(sqrt(u1) * sin(2.0 * M_PI * u3));
  q[3] = (float)(sqrt(u1) * cos(2.0 * M_PI * u3));
}

typedef struct {
  double c[3], s, a;
} blob;

static void make_blobs(const emc_synth_params *p, blob *b) {
  for (int j = 0; j < p->n_blobs; j++) {
    double u[3];
    for (int t = 0;; t++) {
      double r2 = 0;
      for (int k = 0; k < 3; k++) {
        u[k] = 2.0 * emc_uniform(p->seed, ST_BLOB, (uint64_t)j, (uint64_t)(3 * t + k)) - 1.0;
        r2 += u[k] * u[k];
      }
      if (r2 <= 1.0 || t > 1000) break;
    }
    for (int k = 0; k < 3; k++) b[j].c[k] = p->blob_radius * u[k];
    b[j].s = p->sigma_min + (p->sigma_max - p->sigma_min) * emc_uniform(p->seed, ST_BLOB, (uint64_t)j, 100000);
    b[j].a = 0.5 + 0.5 * emc_uniform(p->seed, ST_BLOB, (uint64_t)j, 100001);
  }
}

int emc_synth_intensity(const emc_synth_params *p, float *volume) {
  const int S = p->side, nb = p->n_blobs;
  if (S <= 0 || nb <= 0) return -1;
  blob *b = (blob *)malloc(sizeof(blob) * nb);
  make_blobs(p, b);
  /* per-blob, per-axis complex factors g[j][axis][v] */
  double *gre = (double *)malloc(sizeof(double) * (size_t)nb * 3 * S);
  double *gim = (double *)malloc(sizeof(double) * (size_t)nb * 3 * S);
  for (int j = 0; j < nb; j++) {
    const double amp = b[j].a * pow(2.0 * M_PI * b[j].s * b[j].s, 1.5);
    for (int ax = 0; ax < 3; ax++)
      for (int v = 0; v < S; v++) {
        const double k = v - S / 2.0 - 0.5;
        const double env = exp(-2.0 * M_PI * M_PI * b[j].s * b[j].s * k * k / ((double)S * S));
        const double ph = -2.0 * M_PI * k * b[j].c[ax] / S;
        const double sc = ax == 0 ? amp * env : env; /* amplitude folded into the x factor */
        gre[((size_t)j * 3 + ax) * S + v] = sc * cos(ph);
        gim[((size_t)j * 3 + ax) * S + v] = sc * sin(ph);
      }
  }
  double *slab = (double *)calloc(S, sizeof(double));
#pragma omp parallel
  {
    double *fre = (double *)malloc(sizeof(double) * S);
    double *fim = (double *)malloc(sizeof(double) * S);
#pragma omp for schedule(static)
    for (int z = 0; z < S; z++) {
      double zs = 0.0;
      for (int y = 0; y < S; y++) {
        for (int x = 0; x < S; x++) fre[x] = fim[x] = 0.0;
        for (int j = 0; j < nb; j++) {
          const double *xr = &gre[((size_t)j * 3 + 0) * S], *xi = &gim[((size_t)j * 3 + 0) * S];
          const double yr = gre[((size_t)j * 3 + 1) * S + y], yi = gim[((size_t)j * 3 + 1) * S + y];
          const double zr = gre[((size_t)j * 3 + 2) * S + z], zi = gim[((size_t)j * 3 + 2) * S + z];
          const double wr = yr * zr - yi * zi, wi = yr * zi + yi * zr;
          for (int x = 0; x < S; x++) {
            fre[x] += wr * xr[x] - wi * xi[x];
            fim[x] += wr * xi[x] + wi * xr[x];
          }
        }
        for (int x = 0; x < S; x++) {
          const double I = fre[x] * fre[x] + fim[x] * fim[x];
          volume[((size_t)z * S + y) * S + x] = (float)I;
          zs += I;
        }
      }
      slab[z] = zs;
    }
    free(fre);
    free(fim);
  }
  double tot = 0.0;
  for (int z = 0; z < S; z++) tot += slab[z];
  const double mean = tot / ((double)S * S * S);
  const size_t n3 = (size_t)S * S * S;
#pragma omp parallel for schedule(static)
  for (size_t k = 0; k < n3; k++) volume[k] = (float)(volume[k] / mean);
  free(slab);
  free(gre);
  free(gim);
  free(b);
  return 0;
}

/* Trilinear sample of the volume at continuous voxel coordinate u (zero
 * outside the grid). */
static double trilinear(const float *vol, int S, double ux, double uy, double uz) {
  const double fx0 = floor(ux), fy0 = floor(uy), fz0 = floor(uz);
  const int x0 = (int)fx0, y0 = (int)fy0, z0 = (int)fz0;
  const double fx = ux - fx0, fy = uy - fy0, fz = uz - fz0;
  double acc = 0.0;
  for (int c = 0; c < 8; c++) {
    const int x = x0 + (c & 1), y = y0 + ((c >> 1) & 1), z = z0 + ((c >> 2) & 1);
    if (x < 0 || x >= S || y < 0 || y >= S || z < 0 || z >= S) continue;
    const double w = ((c & 1) ? fx : 1 - fx) * (((c >> 1) & 1) ? fy : 1 - fy) * (((c >> 2) & 1) ? fz : 1 - fz);
    acc += w * vol[((size_t)z * S + y) * S + x];
  }
  return acc;
}

/* Poisson variate from the counter stream (seed, ST_COUNT, key, t), t = 0,1,..
 * Inversion for lam < 12, Hormann's transformed rejection (PTRS) above. */
static int poisson(uint64_t seed, uint64_t key, double lam) {
  if (!(lam > 0.0)) return 0;
  uint64_t t = 0;
  if (lam < 12.0) {
    const double u = emc_uniform(seed, ST_COUNT, key, t);
    double pk = exp(-lam), F = pk;
    int k = 0;
    while (u > F && k < 1000) {
      k++;
      pk *= lam / k;
      F += pk;
    }
    return k;
  }
  const double slam = sqrt(lam), loglam = log(lam);
  const double b = 0.931 + 2.53 * slam, a = -0.059 + 0.02483 * b;
  const double invalpha = 1.1239 + 1.1328 / (b - 3.4), vr = 0.9277 - 3.6224 / (b - 2);
  for (int it = 0; it < 10000; it++) {
    const double U = emc_uniform(seed, ST_COUNT, key, t++) - 0.5;
    const double V = emc_uniform(seed, ST_COUNT, key, t++);
    const double us = 0.5 - fabs(U);
    const double k = floor((2 * a / us + b) * U + lam + 0.43);
    if (us >= 0.07 && V <= vr) return (int)k;
    if (k < 0 || (us < 0.013 && V > us)) continue;
    if (log(V) + log(invalpha) - log(a / (us * us) + b) <= -lam + k * loglam - lgamma(k + 1)) return (int)k;
  }
  return (int)floor(lam + 0.5);
}

int emc_synth_patterns(const emc_synth_params *p, const float *volume, const float *x,
                       const float *y, const float *z, const int *mask, float *images,
                       float *q_true, float *phi_true) {
  const int S = p->side, N = p->n_patterns;
  const size_t N2d = (size_t)S * S;
  if (S <= 0 || N <= 0) return -1;
  float *qv = (float *)malloc(sizeof(float) * 4 * (size_t)N);
  double *phi = (double *)malloc(sizeof(double) * N);
  double *isum = (double *)malloc(sizeof(double) * N);
  for (int i = 0; i < N; i++) {
    emc_quaternion_from_uniforms(emc_uniform(p->seed, ST_ORIENT, (uint64_t)i, 0),
                                 emc_uniform(p->seed, ST_ORIENT, (uint64_t)i, 1),
                                 emc_uniform(p->seed, ST_ORIENT, (uint64_t)i, 2), &qv[4 * i]);
    const double u1 = emc_uniform(p->seed, ST_PHI, (uint64_t)i, 0), u2 = emc_uniform(p->seed, ST_PHI, (uint64_t)i, 1);
    const double g = sqrt(-2.0 * log(1.0 - u1)) * cos(2.0 * M_PI * u2);
    phi[i] = exp(p->phi_sigma * g);
  }
  double phisum = 0.0;
  for (int i = 0; i < N; i++) phisum += phi[i];
  for (int i = 0; i < N; i++) phi[i] /= phisum / N; /* mean 1, as emc.c:793-796 */

  /* pass 1: noiseless intensity of every masked pixel */
#pragma omp parallel for schedule(dynamic, 4)
  for (int i = 0; i < N; i++) {
    const double a = qv[4 * i], b = qv[4 * i + 1], c = qv[4 * i + 2], d = qv[4 * i + 3];
    const double m[9] = {a * a + b * b - c * c - d * d, 2 * b * c - 2 * a * d, 2 * b * d + 2 * a * c,
                         2 * b * c + 2 * a * d, a * a - b * b + c * c - d * d, 2 * c * d - 2 * a * b,
                         2 * b * d - 2 * a * c, 2 * c * d + 2 * a * b, a * a - b * b - c * c + d * d};
    float *img = &images[(size_t)i * N2d];
    double s = 0.0;
    for (size_t px = 0; px < N2d; px++) {
      if (!mask[px]) { img[px] = 0.0f; continue; }
      const double cx = x[px], cy = y[px], cz = z[px];
      const double o = S / 2.0 + 0.5;
      const double v = trilinear(volume, S, o + m[0] * cx + m[1] * cy + m[2] * cz,
                                 o + m[3] * cx + m[4] * cy + m[5] * cz, o + m[6] * cx + m[7] * cy + m[8] * cz);
      img[px] = (float)v;
      s += v;
    }
    isum[i] = s;
  }
  double tot = 0.0;
  for (int i = 0; i < N; i++) tot += phi[i] * isum[i];
  const double scale = tot > 0 ? p->photons / (tot / N) : 0.0;
  /* pass 2: Poisson counts */
#pragma omp parallel for schedule(dynamic, 4)
  for (int i = 0; i < N; i++) {
    float *img = &images[(size_t)i * N2d];
    for (size_t px = 0; px < N2d; px++) {
      if (!mask[px]) continue;
      img[px] = (float)poisson(p->seed, (uint64_t)i * N2d + px, scale * phi[i] * img[px]);
    }
  }
  if (q_true) memcpy(q_true, qv, sizeof(float) * 4 * (size_t)N);
  if (phi_true)
    for (int i = 0; i < N; i++) phi_true[i] = (float)phi[i];
  free(qv);
  free(phi);
  free(isum);
  return 0;
}

void emc_synth_random_model(int side, uint64_t seed, float *model) {
  const size_t n3 = (size_t)side * side * side;
#pragma omp parallel for schedule(static)
  for (size_t k = 0; k < n3; k++) model[k] = (float)emc_uniform(seed, ST_MODEL, k, 0);
}

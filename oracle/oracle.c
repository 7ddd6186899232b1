/*
 * CPU oracle for the EMC iteration -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Compiled with -ffp-contract=off so that every float expression is evaluated
 * exactly as written (no fused multiply-add), with FLT_EVAL_METHOD 0 (x86-64
 * SSE): float op float is rounded to float, float op double to double.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NTHR 256 /* threads per block in every reference kernel */

int o_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void o_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ------------------------------------------------------------------------ */
/* Rotation table: restatement of emc.c:40-363.                              */
/* ------------------------------------------------------------------------ */

/* emc.c:41 -- truncated sqrt(5), evaluated in double, stored as float. */
static const float O_TAU = (float)((1.0 + 2.23606798) / 2.0);

/* emc.c:43 */
int o_n_to_samples(int n) { return (int)(20 * (n + 5 * pow(n, 3))); }

typedef struct { float q[4]; } oq;

/* emc.c:50-103: the 120 vertices of the 600-cell. */
static void o_vertices(oq *v) {
  for (int i = 0; i < 120; i++) { v[i].q[0] = 0.0f; v[i].q[1] = 0.0f; v[i].q[2] = 0.0f; v[i].q[3] = 0.0f; }
  for (int i1 = 0; i1 < 2; i1++)
    for (int i2 = 0; i2 < 2; i2++)
      for (int i3 = 0; i3 < 2; i3++)
        for (int i4 = 0; i4 < 2; i4++) {
          oq *p = &v[8 * i1 + 4 * i2 + 2 * i3 + i4];
          p->q[0] = (float)(-0.5 + (float)i1);
          p->q[1] = (float)(-0.5 + (float)i2);
          p->q[2] = (float)(-0.5 + (float)i3);
          p->q[3] = (float)(-0.5 + (float)i4);
        }
  for (int i = 0; i < 8; i++) v[16 + i].q[i / 2] = (float)(-1.0 + 2.0 * (float)(i % 2));
  static const int it[12][4] = {{1, 2, 3, 4}, {1, 4, 2, 3}, {1, 3, 4, 2}, {2, 3, 1, 4},
                                {2, 4, 3, 1}, {2, 1, 4, 3}, {3, 1, 2, 4}, {3, 4, 1, 2},
                                {3, 2, 4, 1}, {4, 2, 1, 3}, {4, 3, 2, 1}, {4, 1, 3, 2}};
  for (int i = 0; i < 12; i++)
    for (int j1 = 0; j1 < 2; j1++)
      for (int j2 = 0; j2 < 2; j2++)
        for (int j3 = 0; j3 < 2; j3++) {
          oq *p = &v[24 + 8 * i + 4 * j1 + 2 * j2 + j3];
          p->q[it[i][0] - 1] = (float)(-0.5 + 1.0 * (float)j1);
          p->q[it[i][1] - 1] = (float)((double)O_TAU * (-0.5 + 1.0 * (float)j2));
          p->q[it[i][2] - 1] = (float)(1.0 / (double)O_TAU * (-0.5 + 1.0 * (float)j3));
          p->q[it[i][3] - 1] = 0.0f;
        }
}

static float o_sq4(float a, float b, float c, float d) {
  /* dist2 = pow(a,2)+pow(b,2)+pow(c,2)+pow(d,2) in double, stored to float */
  return (float)(pow(a, 2) + pow(b, 2) + pow(c, 2) + pow(d, 2));
}

static float o_dist3(const oq *p) {
  return (float)pow(pow(p->q[0], 2) + pow(p->q[1], 2) + pow(p->q[2], 2) + pow(p->q[3], 2),
                    (double)((float)3 / (float)2));
}

int o_rotation_list(int n, float *qout, float *wout) {
  oq v[120];
  o_vertices(v);
  /* emc.c:105-129 edges: pairs with |qi+qj|^2 > 3 */
  static int edges[720][2];
  int count = 0;
  for (int i = 0; i < 120; i++)
    for (int j = 0; j < i; j++) {
      float d2 = o_sq4(v[i].q[0] + v[j].q[0], v[i].q[1] + v[j].q[1], v[i].q[2] + v[j].q[2],
                       v[i].q[3] + v[j].q[3]);
      if (d2 > 3.0f) {
        if (count >= 720) return -1;
        edges[count][0] = i;
        edges[count][1] = j;
        count++;
      }
    }
  if (count != 720) return -1;
  /* emc.c:131-168 faces */
  static int faces[1200][3];
  int face_done[120];
  memset(face_done, 0, sizeof(face_done));
  count = 0;
  for (int i = 0; i < 720; i++) {
    face_done[edges[i][0]] = 1;
    face_done[edges[i][1]] = 1;
    for (int j = 0; j < 120; j++) {
      if (face_done[j]) continue;
      const oq *a = &v[j], *b = &v[edges[i][0]], *c = &v[edges[i][1]];
      float d2 = o_sq4(a->q[0] + b->q[0] + c->q[0], a->q[1] + b->q[1] + c->q[1],
                       a->q[2] + b->q[2] + c->q[2], a->q[3] + b->q[3] + c->q[3]);
      if (d2 > 7.5f) {
        if (count >= 1200) return -1;
        faces[count][0] = edges[i][0];
        faces[count][1] = edges[i][1];
        faces[count][2] = j;
        count++;
      }
    }
  }
  if (count != 1200) return -1;
  /* emc.c:170-210 cells */
  static int cells[600][4];
  int cell_done[120];
  memset(cell_done, 0, sizeof(cell_done));
  count = 0;
  for (int j = 0; j < 120; j++) {
    cell_done[j] = 1;
    for (int i = 0; i < 1200; i++) {
      if (cell_done[faces[i][0]] || cell_done[faces[i][1]] || cell_done[faces[i][2]]) continue;
      const oq *a = &v[faces[i][0]], *b = &v[faces[i][1]], *c = &v[faces[i][2]], *d = &v[j];
      float d2 = o_sq4(a->q[0] + b->q[0] + c->q[0] + d->q[0], a->q[1] + b->q[1] + c->q[1] + d->q[1],
                       a->q[2] + b->q[2] + c->q[2] + d->q[2], a->q[3] + b->q[3] + c->q[3] + d->q[3]);
      if (d2 > 13.5f) {
        if (count >= 600) return -1;
        cells[count][0] = faces[i][0];
        cells[count][1] = faces[i][1];
        cells[count][2] = faces[i][2];
        cells[count][3] = j;
        count++;
      }
    }
  }
  if (count != 600) return -1;

  /* emc.c:212-217 weight factors */
  float alpha = (float)acos(1.0 / 3.0);
  float f1 = (float)(5.0 * alpha / 2.0 / M_PI);
  float f0 = (float)(20.0 * (3.0 * alpha - M_PI) / 4.0 / M_PI);
  float f2 = 1.0f, f3 = 1.0f;
  int ns = o_n_to_samples(n);
  oq *out = (oq *)calloc((size_t)ns, sizeof(oq));
  float *w = wout;
  /* emc.c:229-240 vertices */
  for (int i = 0; i < 120; i++) {
    out[i] = v[i];
    w[i] = f0 / (float)ns / o_dist3(&out[i]);
  }
  /* emc.c:242-259 edge interior points */
  int ev = n - 1;
  for (int i = 0; i < 720; i++)
    for (int j = 0; j < ev; j++) {
      int idx = 120 + ev * i + j;
      for (int k = 0; k < 4; k++)
        out[idx].q[k] = (float)(j + 1) / (float)(ev + 1) * v[edges[i][0]].q[k] +
                        (float)(ev - j) / (float)(ev + 1) * v[edges[i][1]].q[k];
      w[idx] = f1 / (float)ns / o_dist3(&out[idx]);
    }
  /* emc.c:261-294 face interior points */
  int fbase = 120 + 720 * ev;
  int fv = ((n - 1) * (n - 2)) / 2;
  if (fv > 0) {
    for (int i = 0; i < 1200; i++) {
      int c = 0;
      for (int ka = 2; ka < ev + 1; ka++)
        for (int kb = 2; kb < ev + 1; kb++) {
          if (ka + kb > ev + 1) {
            int kc = 2 * (ev + 1) - ka - kb;
            float den = (float)(3 * (ev + 1) - ka - kb - kc);
            float a = (float)(ev + 1 - ka) / den;
            float b = (float)(ev + 1 - kb) / den;
            float cc = (float)(ev + 1 - kc) / den;
            int idx = fbase + fv * i + c;
            for (int k = 0; k < 4; k++)
              out[idx].q[k] = a * v[faces[i][0]].q[k] + b * v[faces[i][1]].q[k] + cc * v[faces[i][2]].q[k];
            w[idx] = f2 / (float)ns / o_dist3(&out[idx]);
            c++;
          }
        }
    }
  }
  /* emc.c:296-334 cell interior points */
  int cbase = 120 + 720 * ev + 1200 * fv;
  int cv = ((n - 1) * (n - 2) * (n - 3)) / 6;
  if (cv > 0) {
    for (int i = 0; i < 600; i++) {
      int c = 0;
      for (int ka = 3; ka < ev + 1; ka++)
        for (int kb = 3; kb < ev + 1; kb++)
          for (int kc = 3; kc < ev + 1; kc++) {
            int kd = 3 * (ev + 1) - ka - kb - kc;
            if (kd >= 3 && kd < ev + 1) {
              float den = (float)(4 * (ev + 1) - ka - kb - kc - kd);
              float a = (float)(ev + 1 - ka) / den;
              float b = (float)(ev + 1 - kb) / den;
              float cc = (float)(ev + 1 - kc) / den;
              float d = (float)(ev + 1 - kd) / den;
              int idx = cbase + cv * i + c;
              for (int k = 0; k < 4; k++)
                out[idx].q[k] = a * v[cells[i][0]].q[k] + b * v[cells[i][1]].q[k] +
                                cc * v[cells[i][2]].q[k] + d * v[cells[i][3]].q[k];
              w[idx] = f3 / (float)ns / o_dist3(&out[idx]);
              c++;
            }
          }
    }
  }
  /* emc.c:337-345 normalise weights (float running sum) */
  float wsum = 0.0f;
  for (int i = 0; i < ns; i++) wsum += w[i];
  for (int i = 0; i < ns; i++) w[i] /= wsum;
  /* emc.c:355-360 -> emc.c:18-25 normalise quaternions */
  for (int i = 0; i < ns; i++) {
    float ab = (float)sqrt(pow(out[i].q[0], 2) + pow(out[i].q[1], 2) + pow(out[i].q[2], 2) + pow(out[i].q[3], 2));
    for (int k = 0; k < 4; k++) qout[4 * i + k] = out[i].q[k] / ab;
  }
  free(out);
  return ns;
}

/* Antipodal pairing.  q and -q are the same rotation; the table holds both,
 * equal up to float rounding of their construction.  Each entry's partner is
 * found by a tolerance search (1e-6 per component) over the entries sorted by
 * q0, and the first entry of each pair in table order is kept, with weight
 * w(q) + w(-q). */
static const float *o_sort_q;
static int o_cmp_q0(const void *a, const void *b) {
  float x = o_sort_q[4 * *(const int *)a], y = o_sort_q[4 * *(const int *)b];
  if (x < y) return -1;
  if (x > y) return 1;
  return *(const int *)a - *(const int *)b;
}

int o_rotation_list_half(int n, float *qout, float *wout) {
  int ns = o_n_to_samples(n);
  float *q = (float *)malloc(sizeof(float) * 4 * ns);
  float *w = (float *)malloc(sizeof(float) * ns);
  if (o_rotation_list(n, q, w) != ns) { free(q); free(w); return -1; }
  int *ord = (int *)malloc(sizeof(int) * ns);
  for (int i = 0; i < ns; i++) ord[i] = i;
  o_sort_q = q;
  qsort(ord, ns, sizeof(int), o_cmp_q0);
  int *partner = (int *)malloc(sizeof(int) * ns);
  const float tol = 1e-6f;
  for (int i = 0; i < ns; i++) {
    partner[i] = -1;
    float t = -q[4 * i];
    int lo = 0, hi = ns;
    while (lo < hi) {
      int mid = (lo + hi) / 2;
      if (q[4 * ord[mid]] < t - tol) lo = mid + 1; else hi = mid;
    }
    for (int j = lo; j < ns && q[4 * ord[j]] <= t + tol; j++) {
      int c = ord[j];
      int ok = 1;
      for (int k = 0; k < 4; k++) if (fabsf(q[4 * c + k] + q[4 * i + k]) > tol) ok = 0;
      if (ok) { partner[i] = c; break; }
    }
  }
  int *kept = (int *)calloc(ns, sizeof(int));
  int m = 0;
  for (int i = 0; i < ns; i++) {
    int p = partner[i];
    if (p < 0 || p == i) { m = -1; break; }
    if (kept[p]) continue; /* partner already represents this rotation */
    kept[i] = 1;
    for (int k = 0; k < 4; k++) qout[4 * m + k] = q[4 * i + k];
    wout[m] = (float)((double)w[i] + (double)w[p]);
    m++;
  }
  free(kept); free(partner); free(ord); free(q); free(w);
  return m;
}

/* emc.c:365-393, layout [y*side + x]. */
void o_coordinates(int side, float wavelength, float pixel_size, float detector_distance,
                   float *xo, float *yo, float *zo) {
  const int x_max = side, y_max = side;
  for (int x = 0; x < x_max; x++)
    for (int y = 0; y < y_max; y++) {
      float pixel_r = (float)sqrt(pow((float)(x - x_max / 2) + 0.5, 2) + pow((float)(y - y_max / 2) + 0.5, 2));
      float real_r = pixel_r * pixel_size;
      float angle_r = (float)atan2(real_r, detector_distance);
      float fourier_r = (float)(sin(angle_r) / wavelength);
      float fourier_z = (float)((1. - cos(angle_r)) / wavelength);
      float px = (float)((float)(x - x_max / 2) + 0.5);
      float py = (float)((float)(y - y_max / 2) + 0.5);
      float pz = fourier_z / fourier_r * pixel_r;
      xo[y * side + x] = px;
      yo[y * side + x] = py;
      zo[y * side + x] = pz;
    }
}

/* Rotation matrix entries exactly as emc_cuda.cu:65-85 writes them in float. */
static void o_rotmat(const float *r, float e[9]) {
  e[0] = r[0] * r[0] + r[1] * r[1] - r[2] * r[2] - r[3] * r[3];
  e[1] = 2.0f * r[1] * r[2] - 2.0f * r[0] * r[3];
  e[2] = 2.0f * r[1] * r[3] + 2.0f * r[0] * r[2];
  e[3] = 2.0f * r[1] * r[2] + 2.0f * r[0] * r[3];
  e[4] = r[0] * r[0] - r[1] * r[1] + r[2] * r[2] - r[3] * r[3];
  e[5] = 2.0f * r[2] * r[3] - 2.0f * r[0] * r[1];
  e[6] = 2.0f * r[1] * r[3] - 2.0f * r[0] * r[2];
  e[7] = 2.0f * r[2] * r[3] + 2.0f * r[0] * r[1];
  e[8] = r[0] * r[0] - r[1] * r[1] - r[2] * r[2] + r[3] * r[3];
}

static void o_rotate(const float e[9], float xc, float yc, float zc, float *nx, float *ny, float *nz) {
  *nx = e[0] * xc + e[1] * yc + e[2] * zc;
  *ny = e[3] * xc + e[4] * yc + e[5] * zc;
  *nz = e[6] * xc + e[7] * yc + e[8] * zc;
}

void o_voxel_index(const float *q, float xc, float yc, float zc, int S, int *vx, int *vy, int *vz) {
  float e[9], nx, ny, nz;
  o_rotmat(q, e);
  o_rotate(e, xc, yc, zc, &nx, &ny, &nz);
  /* emc_cuda.cu:86-88 */
  *vx = (int)roundf((float)S / 2.0f + 0.5f + nx);
  *vy = (int)roundf((float)S / 2.0f + 0.5f + ny);
  *vz = (int)roundf((float)S / 2.0f + 0.5f + nz);
}

/* ------------------------------------------------------------------------ */
/* FP32 restatement of the reference kernels                                 */
/* ------------------------------------------------------------------------ */

/* emc_cuda.cu:51-111 -- bounds 0 < v < S. */
void r32_expand(const float *model, int S, const float *q, int R, const float *x,
                const float *y, const float *z, float *slices) {
  const size_t N2d = (size_t)S * S;
#pragma omp parallel for schedule(static)
  for (int r = 0; r < R; r++) {
    float e[9];
    o_rotmat(&q[4 * r], e);
    float *sl = &slices[(size_t)r * N2d];
    for (size_t p = 0; p < N2d; p++) {
      float nx, ny, nz;
      o_rotate(e, x[p], y[p], z[p], &nx, &ny, &nz);
      int vx = (int)roundf((float)S / 2.0f + 0.5f + nx);
      int vy = (int)roundf((float)S / 2.0f + 0.5f + ny);
      int vz = (int)roundf((float)S / 2.0f + 0.5f + nz);
      if (vx > 0 && vx < S && vy > 0 && vy < S && vz > 0 && vz < S)
        sl[p] = model[(size_t)vz * S * S + (size_t)vy * S + vx];
      else
        sl[p] = 0.0f;
    }
  }
}

/* Block tree reduction (emc_cuda.cu:25-34) over NTHR partials. */
static float o_tree_sum(float *d) {
  for (int s = NTHR / 2; s > 0; s >>= 1)
    for (int t = 0; t < s; t++) d[t] += d[t + s];
  return d[0];
}
static int o_tree_isum(int *d) {
  for (int s = NTHR / 2; s > 0; s >>= 1)
    for (int t = 0; t < s; t++) d[t] += d[t + s];
  return d[0];
}
/* emc_cuda.cu:36-47 */
static float o_tree_max(float *d) {
  for (int s = NTHR / 2; s > 0; s >>= 1)
    for (int t = 0; t < s; t++)
      if (d[t] < d[t + s]) d[t] = d[t + s];
  return d[0];
}

/* emc_cuda.cu:113-150: per (pattern, rotation) block of 256 threads, thread t
 * sums pixels p = t, t+256, ... in float; block tree reduction; final scale
 * in double.  Pixel p contributes (S_rp - K_ip/phi_i)^2 when mask != 0. */
void r32_likelihood(const float *slices, const float *images, const int *mask, float sigma,
                    const float *scaling, float *respons, int N2d, int N, int R) {
  const double sig2 = (double)sigma * (double)sigma;
  int cnt[NTHR];
  for (int t = 0; t < NTHR; t++) cnt[t] = 0;
  for (int p = 0; p < N2d; p++) cnt[p % NTHR] += mask[p] != 0;
  const int count = o_tree_isum(cnt);
#pragma omp parallel
  {
    float *kd = (float *)malloc(sizeof(float) * (size_t)N2d);
    float part[NTHR];
#pragma omp for schedule(dynamic, 1)
    for (int i = 0; i < N; i++) {
      const float *img = &images[(size_t)i * N2d];
      const float sc = scaling[i];
      for (int p = 0; p < N2d; p++) kd[p] = img[p] / sc;
      for (int r = 0; r < R; r++) {
        const float *sl = &slices[(size_t)r * N2d];
        for (int t = 0; t < NTHR; t++) part[t] = 0.0f;
        int base = 0;
        for (; base + NTHR <= N2d; base += NTHR)
          for (int t = 0; t < NTHR; t++) {
            float d = sl[base + t] - kd[base + t];
            part[t] += mask[base + t] != 0 ? d * d : 0.0f;
          }
        for (int t = 0; base + t < N2d; t++) {
          float d = sl[base + t] - kd[base + t];
          part[t] += mask[base + t] != 0 ? d * d : 0.0f;
        }
        float sum = o_tree_sum(part);
        respons[(size_t)r * N + i] = (float)(-(double)sum / 2.0 / (double)(float)count / sig2);
      }
    }
    free(kd);
  }
}

/* emc_cuda.cu:445-476, with the block maximum taken before any thread clears
 * the reduction buffer (the intended semantics). */
void r32_softmax(float *respons, int R, int N) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; i++) {
    float cache[NTHR];
    for (int t = 0; t < NTHR; t++) cache[t] = -1.0e10f;
    for (int r = 0; r < R; r++) {
      float v = respons[(size_t)r * N + i];
      if (cache[r % NTHR] < v) cache[r % NTHR] = v;
    }
    float mx = o_tree_max(cache);
    for (int t = 0; t < NTHR; t++) cache[t] = 0.0f;
    for (int r = 0; r < R; r++) {
      float v = respons[(size_t)r * N + i] - mx;
      if (v > -1.0e10f) {
        v = expf(v);
        cache[r % NTHR] += v;
      } else {
        v = 0.0f;
      }
      respons[(size_t)r * N + i] = v;
    }
    float sum = o_tree_sum(cache);
    for (int r = 0; r < R; r++) respons[(size_t)r * N + i] /= sum;
  }
}

/* emc_cuda.cu:484-505.  thrust::transform_reduce's order is not specified;
 * terms are float x*logf(x), accumulated in double here. */
double r32_total_respons(const float *respons, size_t n) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (size_t k = 0; k < n; k++) {
    float x = respons[k];
    s += x > 0 ? (double)(x * logf(x)) : 0.0;
  }
  return s;
}

/* emc_cuda.cu:195-228 */
void r32_update_scaling(const float *images, const float *slices, const int *mask,
                        const float *respons, float *scaling, int N, int R, int N2d) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int i = 0; i < N; i++) {
    const float *img = &images[(size_t)i * N2d];
    float part[NTHR];
    for (int t = 0; t < NTHR; t++) part[t] = 0.0f;
    for (int p = 0; p < N2d; p++)
      if (mask[p] != 0) part[p % NTHR] += img[p] * img[p];
    float power = o_tree_sum(part);
    float weighted = 0.0f;
    for (int r = 0; r < R; r++) {
      const float *sl = &slices[(size_t)r * N2d];
      for (int t = 0; t < NTHR; t++) part[t] = 0.0f;
      int base = 0;
      for (; base + NTHR <= N2d; base += NTHR)
        for (int t = 0; t < NTHR; t++)
          part[t] += mask[base + t] != 0 ? img[base + t] * sl[base + t] : 0.0f;
      for (int t = 0; base + t < N2d; t++)
        part[t] += mask[base + t] != 0 ? img[base + t] * sl[base + t] : 0.0f;
      float corr = o_tree_sum(part);
      weighted += respons[(size_t)r * N + i] * corr;
    }
    scaling[i] = power / weighted;
  }
}

/* emc_atomic.cu:17-73 for one slice, sequential (x outer, y inner). */
void r32_insert_slice(float *model, float *weight, int S, const float *slice, const int *mask,
                      float w, const float *q, const float *x, const float *y, const float *z) {
  float e[9];
  o_rotmat(q, e);
  for (int xi = 0; xi < S; xi++)
    for (int yi = 0; yi < S; yi++) {
      int p = yi * S + xi;
      if (mask[p] != 1) continue;
      float nx, ny, nz;
      o_rotate(e, x[p], y[p], z[p], &nx, &ny, &nz);
      int vx = (int)roundf((float)S / 2.0f + 0.5f + nx);
      int vy = (int)roundf((float)S / 2.0f + 0.5f + ny);
      int vz = (int)roundf((float)S / 2.0f + 0.5f + nz);
      if (vx >= 0 && vx < S && vy >= 0 && vy < S && vz >= 0 && vz < S) {
        size_t v = (size_t)vz * S * S + (size_t)vy * S + vx;
        model[v] += w * slice[p];
        weight[v] += w;
      }
    }
}

/* emc_atomic.cu:76-113: the M-step alone (slices overwritten at masked
 * pixels, T written). */
void r32_mstep(const float *images, float *slices, const int *mask, const float *respons,
               const float *scaling, int N, int R, int N2d, float *T) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int r = 0; r < R; r++) {
    const float *pr = &respons[(size_t)r * N];
    float *sl = &slices[(size_t)r * N2d];
    for (int p = 0; p < N2d; p++) {
      if (mask[p] == 0) continue;
      float sum = 0.0f;
      for (int i = 0; i < N; i++) sum += images[(size_t)i * N2d + p] * pr[i] / scaling[i];
      sl[p] = sum;
    }
    float tot = 0.0f;
    for (int i = 0; i < N; i++) tot += pr[i];
    T[r] = tot;
    if (tot > 1e-10f)
      for (int p = 0; p < N2d; p++)
        if (mask[p] != 0) sl[p] /= tot;
  }
}

/* emc_atomic.cu:115-133: insert every slice with T_r > 1e-10, weight
 * weights[r] * T_r.  Float atomics in the reference have no fixed order;
 * here contributions are added in rotation order, sequentially. */
void r32_insert_all(float *model, float *weight, int S, const float *slices, const int *mask,
                    const float *T, const float *q, const float *weights, int R, const float *x,
                    const float *y, const float *z) {
  const size_t N2d = (size_t)S * S;
  for (int r = 0; r < R; r++)
    if (T[r] > 1e-10f)
      r32_insert_slice(model, weight, S, &slices[(size_t)r * N2d], mask, weights[r] * T[r], &q[4 * r], x, y, z);
}

/* The same insertion spread over host threads (rotations split between
 * threads, thread-private accumulators summed at the end): used to time the
 * reference's method on all host cores.  Summation order differs from
 * r32_insert_all, as the reference's atomics' order does. */
void r32_insert_all_par(float *model, float *weight, int S, const float *slices, const int *mask,
                        const float *T, const float *q, const float *weights, int R, const float *x,
                        const float *y, const float *z) {
  const size_t N2d = (size_t)S * S, n3 = (size_t)S * S * S;
#pragma omp parallel
  {
    float *m = (float *)calloc(n3, sizeof(float));
    float *w = (float *)calloc(n3, sizeof(float));
#pragma omp for schedule(dynamic, 16)
    for (int r = 0; r < R; r++)
      if (T[r] > 1e-10f)
        r32_insert_slice(m, w, S, &slices[(size_t)r * N2d], mask, weights[r] * T[r], &q[4 * r], x, y, z);
#pragma omp critical
    for (size_t k = 0; k < n3; k++) {
      model[k] += m[k];
      weight[k] += w[k];
    }
    free(m);
    free(w);
  }
}

/* emc_atomic.cu:76-133 and emc_cuda.cu:348-351 (M-step, insert, sum T). */
float r32_update_slices(const float *images, float *slices, const int *mask, const float *respons,
                        const float *scaling, int N, int R, int N2d, int S, float *model,
                        float *weight, const float *x, const float *y, const float *z,
                        const float *q, const float *weights, float *T) {
  r32_mstep(images, slices, mask, respons, scaling, N, R, N2d, T);
  r32_insert_all(model, weight, S, slices, mask, T, q, weights, R, x, y, z);
  float overall = 0.0f; /* emc_cuda.cu:348-351 */
  for (int r = 0; r < R; r++) overall += T[r];
  return overall;
}

/* emc_cuda.cu:414-433.  thrust::reduce order unspecified: double sum here. */
void r32_normalize_model(float *model, const float *weight, int n) {
  for (int k = 0; k < n; k++) model[k] = weight[k] > 0.0f ? model[k] / weight[k] : 0.0f;
  double s = 0.0;
  for (int k = 0; k < n; k++) s += model[k];
  float msum = (float)s;
  msum /= n;
  float inv = 1.0f / msum;
  for (int k = 0; k < n; k++) model[k] *= inv;
}

/* emc_cuda.cu:364-368 */
float r32_model_max(const float *model, int n) {
  float m = 0.0f;
  for (int k = 0; k < n; k++) m = model[k] > m ? model[k] : m;
  return m;
}

/* emc.c:989-1084 step order (expand, likelihood, softmax, total, reset,
 * scaling, M-step + insert, normalise). */
double r32_iteration(float *model, float *weight, float *slices, const float *images,
                     const int *mask, float *scaling, float *respons, int N, int R, int S,
                     const float *x, const float *y, const float *z, const float *q,
                     const float *weights, float sigma, int nthreads) {
#ifdef _OPENMP
  int saved = omp_get_max_threads();
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  const int N2d = S * S;
  const int n3 = S * S * S;
  r32_expand(model, S, q, R, x, y, z, slices);
  r32_likelihood(slices, images, mask, sigma, scaling, respons, N2d, N, R);
  r32_softmax(respons, R, N);
  double tot = r32_total_respons(respons, (size_t)N * R);
  memset(model, 0, sizeof(float) * n3);
  memset(weight, 0, sizeof(float) * n3);
  r32_update_scaling(images, slices, mask, respons, scaling, N, R, N2d);
  float *T = (float *)malloc(sizeof(float) * R);
  r32_update_slices(images, slices, mask, respons, scaling, N, R, N2d, S, model, weight, x, y, z,
                    q, weights, T);
  free(T);
  r32_normalize_model(model, weight, n3);
#ifdef _OPENMP
  omp_set_num_threads(saved);
#endif
  return tot;
}

/* ------------------------------------------------------------------------ */
/* FP64 oracle                                                               */
/* ------------------------------------------------------------------------ */

static void o64_rotmat(const float *r, double e[9]) {
  double a = r[0], b = r[1], c = r[2], d = r[3];
  e[0] = a * a + b * b - c * c - d * d;
  e[1] = 2 * b * c - 2 * a * d;
  e[2] = 2 * b * d + 2 * a * c;
  e[3] = 2 * b * c + 2 * a * d;
  e[4] = a * a - b * b + c * c - d * d;
  e[5] = 2 * c * d - 2 * a * b;
  e[6] = 2 * b * d - 2 * a * c;
  e[7] = 2 * c * d + 2 * a * b;
  e[8] = a * a - b * b - c * c + d * d;
}

/* Trilinear footprint of pixel p under rotation q: up to 8 voxels + weights.
 * Continuous voxel coordinate u = S/2 + 0.5 + R c (the same origin as the
 * nearest-neighbour rule); corners outside 0 <= v < S are dropped. */
static int o64_tri(const float *q, float xc, float yc, float zc, int S, size_t *vox, double *wt) {
  double e[9];
  o64_rotmat(q, e);
  double u[3];
  u[0] = S / 2.0 + 0.5 + (e[0] * xc + e[1] * yc + e[2] * zc);
  u[1] = S / 2.0 + 0.5 + (e[3] * xc + e[4] * yc + e[5] * zc);
  u[2] = S / 2.0 + 0.5 + (e[6] * xc + e[7] * yc + e[8] * zc);
  int i0[3];
  double f[3];
  for (int k = 0; k < 3; k++) {
    double fl = floor(u[k]);
    i0[k] = (int)fl;
    f[k] = u[k] - fl;
  }
  int m = 0;
  for (int c = 0; c < 8; c++) {
    int ix = i0[0] + (c & 1), iy = i0[1] + ((c >> 1) & 1), iz = i0[2] + ((c >> 2) & 1);
    double w = ((c & 1) ? f[0] : 1 - f[0]) * (((c >> 1) & 1) ? f[1] : 1 - f[1]) *
               (((c >> 2) & 1) ? f[2] : 1 - f[2]);
    if (ix < 0 || ix >= S || iy < 0 || iy >= S || iz < 0 || iz >= S) continue;
    vox[m] = (size_t)iz * S * S + (size_t)iy * S + ix;
    wt[m] = w;
    m++;
  }
  return m;
}

void o64_expand(const double *model, int S, const float *q, int R, const float *x,
                const float *y, const float *z, int interp, double *slices) {
  const size_t N2d = (size_t)S * S;
#pragma omp parallel for schedule(static)
  for (int r = 0; r < R; r++) {
    for (size_t p = 0; p < N2d; p++) {
      double v = 0.0;
      if (interp == O_NEAREST) {
        int vx, vy, vz;
        o_voxel_index(&q[4 * r], x[p], y[p], z[p], S, &vx, &vy, &vz);
        if (vx > 0 && vx < S && vy > 0 && vy < S && vz > 0 && vz < S)
          v = model[(size_t)vz * S * S + (size_t)vy * S + vx];
      } else {
        size_t vox[8];
        double wt[8];
        int m = o64_tri(&q[4 * r], x[p], y[p], z[p], S, vox, wt);
        for (int k = 0; k < m; k++) v += wt[k] * model[vox[k]];
      }
      slices[(size_t)r * N2d + p] = v;
    }
  }
}

void o64_likelihood(const double *slices, const float *images, const int *mask, double sigma,
                    const double *scaling, int N2d, int N, int R, int kind, double *L) {
  int count = 0;
  for (int p = 0; p < N2d; p++) count += mask[p] != 0;
#pragma omp parallel for schedule(dynamic, 1)
  for (int i = 0; i < N; i++) {
    const float *img = &images[(size_t)i * N2d];
    for (int r = 0; r < R; r++) {
      const double *sl = &slices[(size_t)r * N2d];
      double s = 0.0;
      if (kind == O_GAUSSIAN) {
        for (int p = 0; p < N2d; p++)
          if (mask[p] != 0) {
            double d = sl[p] - (double)img[p] / scaling[i];
            s += d * d;
          }
        L[(size_t)r * N + i] = -s / 2.0 / (double)count / (sigma * sigma);
      } else {
        for (int p = 0; p < N2d; p++)
          if (mask[p] != 0) {
            /* log(phi max(S, 1e-6)) - phi S: the floor is on the slice value */
            const double sv = sl[p] > 1e-6 ? sl[p] : 1e-6;
            s += (double)img[p] * (log(scaling[i]) + log(sv)) - scaling[i] * sl[p];
          }
        L[(size_t)r * N + i] = s;
      }
    }
  }
}

void o64_softmax(const double *L, int R, int N, double *P) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; i++) {
    double m = -1e10;
    for (int r = 0; r < R; r++) m = L[(size_t)r * N + i] > m ? L[(size_t)r * N + i] : m;
    double s = 0.0;
    for (int r = 0; r < R; r++) {
      double d = L[(size_t)r * N + i] - m;
      double e = d > -1e10 ? exp(d) : 0.0;
      P[(size_t)r * N + i] = e;
      s += e;
    }
    for (int r = 0; r < R; r++) P[(size_t)r * N + i] /= s;
  }
}

double o64_total_respons(const double *P, size_t n) {
  double s = 0.0;
  for (size_t k = 0; k < n; k++) s += P[k] > 0 ? P[k] * log(P[k]) : 0.0;
  return s;
}

void o64_update_scaling(const float *images, const double *slices, const int *mask,
                        const double *P, int N, int R, int N2d, int kind, double *scaling) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int i = 0; i < N; i++) {
    const float *img = &images[(size_t)i * N2d];
    double num = 0.0, den = 0.0;
    for (int p = 0; p < N2d; p++)
      if (mask[p] != 0) num += kind == O_GAUSSIAN ? (double)img[p] * img[p] : (double)img[p];
    for (int r = 0; r < R; r++) {
      const double *sl = &slices[(size_t)r * N2d];
      double c = 0.0;
      for (int p = 0; p < N2d; p++)
        if (mask[p] != 0) c += kind == O_GAUSSIAN ? (double)img[p] * sl[p] : sl[p];
      den += P[(size_t)r * N + i] * c;
    }
    scaling[i] = num / den;
  }
}

void o64_mstep(const float *images, const int *mask, const double *P, const double *scaling,
               int N, int R, int N2d, int kind, const double *slices_in, double *slices, double *T) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int r = 0; r < R; r++) {
    const double *pr = &P[(size_t)r * N];
    double t = 0.0, tphi = 0.0;
    for (int i = 0; i < N; i++) { t += pr[i]; tphi += pr[i] * scaling[i]; }
    T[r] = t;
    for (int p = 0; p < N2d; p++) {
      size_t k = (size_t)r * N2d + p;
      if (mask[p] == 0) { slices[k] = slices_in ? slices_in[k] : 0.0; continue; }
      double u = 0.0;
      if (kind == O_GAUSSIAN) {
        for (int i = 0; i < N; i++) u += (double)images[(size_t)i * N2d + p] * pr[i] / scaling[i];
        slices[k] = t > 1e-10 ? u / t : u;
      } else {
        for (int i = 0; i < N; i++) u += (double)images[(size_t)i * N2d + p] * pr[i];
        slices[k] = tphi > 1e-30 ? u / tphi : 0.0;
      }
    }
  }
}

void o64_insert(const double *slices, const int *mask, const double *T, const float *weights,
                int R, int N2d, int S, const float *q, const float *x, const float *y,
                const float *z, int interp, double *model, double *weight) {
  for (int r = 0; r < R; r++) {
    if (!(T[r] > 1e-10)) continue;
    double w = (double)weights[r] * T[r];
    for (int p = 0; p < N2d; p++) {
      if (mask[p] != 1) continue;
      double s = slices[(size_t)r * N2d + p];
      if (interp == O_NEAREST) {
        int vx, vy, vz;
        o_voxel_index(&q[4 * r], x[p], y[p], z[p], S, &vx, &vy, &vz);
        if (vx >= 0 && vx < S && vy >= 0 && vy < S && vz >= 0 && vz < S) {
          size_t v = (size_t)vz * S * S + (size_t)vy * S + vx;
          model[v] += w * s;
          weight[v] += w;
        }
      } else {
        size_t vox[8];
        double wt[8];
        int m = o64_tri(&q[4 * r], x[p], y[p], z[p], S, vox, wt);
        for (int k = 0; k < m; k++) {
          model[vox[k]] += wt[k] * w * s;
          weight[vox[k]] += wt[k] * w;
        }
      }
    }
  }
}

void o64_normalize_model(double *model, const double *weight, int n) {
  double s = 0.0;
  for (int k = 0; k < n; k++) {
    model[k] = weight[k] > 0 ? model[k] / weight[k] : 0.0;
    s += model[k];
  }
  s /= n;
  for (int k = 0; k < n; k++) model[k] /= s;
}

/* ------------------------------------------------------------------------ */
/* FP64 iteration for problems too large for the dense loops above: the same  */
/* equations as o64_expand .. o64_normalize_model (nearest voxel), evaluated  */
/* over the masked pixels only, with every pattern's nonzero counts as a list */
/* and the zero responsibilities skipped in the scaling and M-step sums.     */
/* The Gaussian likelihood uses the expansion                                 */
/*   sum m (S - K/phi)^2 = sum m S^2 - 2 sum m K S / phi + sum m K^2 / phi^2, */
/* exact in real arithmetic; in FP64 its cancellation costs ~1e-12 relative.  */
/* Pinned against the dense o64_* functions on small problems by the tests.   */
/* allowed ([R][N], may be NULL): the softmax of pattern i runs over the r    */
/* with allowed[r*N+i] != 0 only (o64_restricted_softmax's rule).             */
/* ------------------------------------------------------------------------ */
double o64_iteration_large(const double *model_in, const float *images, const int *mask,
                           const double *scaling_in, int N, int R, int S, const float *q,
                           const float *weights, const float *x, const float *y, const float *z,
                           double sigma, int kind, const unsigned char *allowed, double *L,
                           double *P, double *scaling_out, double *model_out, double *raw_out,
                           double *weight_out) {
  const int N2d = S * S;
  const size_t n3 = (size_t)S * S * S;
  int nm = 0;
  for (int p = 0; p < N2d; p++) nm += mask[p] != 0;
  int *pix = (int *)malloc(sizeof(int) * (nm > 0 ? nm : 1));
  nm = 0;
  for (int p = 0; p < N2d; p++)
    if (mask[p] != 0) pix[nm++] = p;
  /* nonzero counts of each pattern over the masked pixels (compact index) */
  size_t *off = (size_t *)malloc(sizeof(size_t) * (N + 1));
  off[0] = 0;
  for (int i = 0; i < N; i++) {
    size_t c = 0;
    for (int k = 0; k < nm; k++) c += images[(size_t)i * N2d + pix[k]] != 0.0f;
    off[i + 1] = off[i] + c;
  }
  int *nzk = (int *)malloc(sizeof(int) * (off[N] > 0 ? off[N] : 1));
  double *nzv = (double *)malloc(sizeof(double) * (off[N] > 0 ? off[N] : 1));
  double *B = (double *)calloc(N, sizeof(double)); /* Gaussian: sum K^2; Poisson: sum K */
  for (int i = 0; i < N; i++) {
    size_t c = off[i];
    for (int k = 0; k < nm; k++) {
      const double v = images[(size_t)i * N2d + pix[k]];
      if (v != 0.0) {
        nzk[c] = k;
        nzv[c++] = v;
        B[i] += kind == O_GAUSSIAN ? v * v : v;
      }
    }
  }
  double *A = (double *)calloc(R, sizeof(double)); /* Gaussian: sum S^2; Poisson: sum S */
  double *Cx = (double *)malloc(sizeof(double) * (size_t)R * N); /* sum m K S (Gaussian) */
  const double inv = 1.0 / (2.0 * (double)nm * sigma * sigma);
  /* expand + likelihood, one rotation per task */
#pragma omp parallel
  {
    double *sl = (double *)malloc(sizeof(double) * (nm > 0 ? nm : 1));
    double *op = (double *)malloc(sizeof(double) * (nm > 0 ? nm : 1));
#pragma omp for schedule(dynamic, 4)
    for (int r = 0; r < R; r++) {
      double a = 0.0;
      for (int k = 0; k < nm; k++) {
        const int p = pix[k];
        int vx, vy, vz;
        o_voxel_index(&q[4 * r], x[p], y[p], z[p], S, &vx, &vy, &vz);
        double v = 0.0;
        if (vx > 0 && vx < S && vy > 0 && vy < S && vz > 0 && vz < S)
          v = model_in[(size_t)vz * S * S + (size_t)vy * S + vx];
        sl[k] = v;
        a += kind == O_GAUSSIAN ? v * v : v;
        op[k] = kind == O_GAUSSIAN ? v : log(v > 1e-6 ? v : 1e-6);
      }
      A[r] = a;
      for (int i = 0; i < N; i++) {
        double c = 0.0;
        for (size_t e = off[i]; e < off[i + 1]; e++) c += nzv[e] * op[nzk[e]];
        const double phi = scaling_in[i];
        Cx[(size_t)r * N + i] = c;
        L[(size_t)r * N + i] = kind == O_GAUSSIAN ? -(a - 2.0 * c / phi + B[i] / (phi * phi)) * inv
                                                  : c + B[i] * log(phi) - phi * a;
      }
    }
    free(sl);
    free(op);
  }
  /* responsibilities: per-pattern softmax (restricted to the allowed set) */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; i++) {
    double m = -1e10;
    for (int r = 0; r < R; r++)
      if (!allowed || allowed[(size_t)r * N + i]) m = L[(size_t)r * N + i] > m ? L[(size_t)r * N + i] : m;
    double s = 0.0;
    for (int r = 0; r < R; r++) {
      const double d = L[(size_t)r * N + i] - m;
      const double e = (!allowed || allowed[(size_t)r * N + i]) && d > -1e10 ? exp(d) : 0.0;
      P[(size_t)r * N + i] = e;
      s += e;
    }
    for (int r = 0; r < R; r++) P[(size_t)r * N + i] /= s;
  }
  double tot = 0.0;
  for (size_t k = 0; k < (size_t)R * N; k++) tot += P[k] > 0 ? P[k] * log(P[k]) : 0.0;
  /* scaling: Gaussian sum K^2 / sum_r P sum K S; Poisson sum K / sum_r P sum S */
  for (int i = 0; i < N; i++) {
    double den = 0.0;
    for (int r = 0; r < R; r++) {
      const double pr = P[(size_t)r * N + i];
      if (pr != 0.0) den += pr * (kind == O_GAUSSIAN ? Cx[(size_t)r * N + i] : A[r]);
    }
    scaling_out[i] = B[i] / den;
  }
  /* M-step and insertion (nearest voxel, mask == 1, T_r > 1e-10): rotations
   * in chunks; each chunk's updated slices and voxel indices are computed in
   * parallel, then added to the model sequentially in rotation order */
  enum { CH = 128 };
  double *sv = (double *)malloc(sizeof(double) * (size_t)CH * (nm > 0 ? nm : 1));
  long *vi = (long *)malloc(sizeof(long) * (size_t)CH * (nm > 0 ? nm : 1));
  double *wr = (double *)malloc(sizeof(double) * CH);
  double *wsum = (double *)calloc(n3, sizeof(double));
  memset(model_out, 0, sizeof(double) * n3);
  for (int r0 = 0; r0 < R; r0 += CH) {
    const int nr = R - r0 < CH ? R - r0 : CH;
#pragma omp parallel for schedule(dynamic, 1)
    for (int j = 0; j < nr; j++) {
      const int r = r0 + j;
      double *u = sv + (size_t)j * nm;
      long *vv = vi + (size_t)j * nm;
      double t = 0.0, tphi = 0.0;
      for (int k = 0; k < nm; k++) u[k] = 0.0;
      for (int i = 0; i < N; i++) {
        const double pr = P[(size_t)r * N + i];
        if (pr == 0.0) continue;
        t += pr;
        tphi += pr * scaling_out[i];
        const double f = kind == O_GAUSSIAN ? pr / scaling_out[i] : pr;
        for (size_t e = off[i]; e < off[i + 1]; e++) u[nzk[e]] += nzv[e] * f;
      }
      wr[j] = t > 1e-10 ? (double)weights[r] * t : 0.0;
      if (!(t > 1e-10)) continue;
      for (int k = 0; k < nm; k++) {
        const int p = pix[k];
        u[k] = kind == O_GAUSSIAN ? u[k] / t : (tphi > 1e-30 ? u[k] / tphi : 0.0);
        int vx, vy, vz;
        o_voxel_index(&q[4 * r], x[p], y[p], z[p], S, &vx, &vy, &vz);
        vv[k] = (mask[p] == 1 && vx >= 0 && vx < S && vy >= 0 && vy < S && vz >= 0 && vz < S)
                    ? (long)vz * S * S + (long)vy * S + vx
                    : -1;
      }
    }
    for (int j = 0; j < nr; j++) {
      if (wr[j] == 0.0) continue;
      for (int k = 0; k < nm; k++) {
        const long v = vi[(size_t)j * nm + k];
        if (v < 0) continue;
        model_out[v] += wr[j] * sv[(size_t)j * nm + k];
        wsum[v] += wr[j];
      }
    }
  }
  if (raw_out) memcpy(raw_out, model_out, sizeof(double) * n3);
  if (weight_out) memcpy(weight_out, wsum, sizeof(double) * n3);
  o64_normalize_model(model_out, wsum, (int)n3);
  free(wsum);
  free(wr);
  free(vi);
  free(sv);
  free(Cx);
  free(A);
  free(B);
  free(nzv);
  free(nzk);
  free(off);
  free(pix);
  return tot;
}
